"""Mole-lite reproduces PySCF's basis conventions."""
import numpy

from pyscf_amd import gto
from tests.conftest import H2O


def test_shell_structure_ccpvdz(h2o_vdz):
    # O: [S 8p x 2c, S, P 3p, P, D]; H: [S 3p, S, P]
    bas = h2o_vdz._bas
    assert h2o_vdz.nbas == 11 and h2o_vdz.nao_nr() == 24
    assert list(bas[:5, 1]) == [0, 0, 1, 1, 2]
    assert list(bas[:5, 2]) == [8, 1, 3, 1, 1] and bas[0, 3] == 2
    assert list(bas[5:8, 1]) == [0, 0, 1]


def test_sp_split_631g(h2o_631g):
    assert h2o_631g.nao_nr() == 13
    assert list(h2o_631g._bas[:5, 1]) == [0, 0, 0, 1, 1]


def test_sizes_of_benchmark_bases():
    assert gto.M(atom=H2O, basis='def2-svp').nao_nr() == 24
    assert gto.M(atom=H2O, basis='def2-tzvp').nao_nr() == 43


def test_primitives_sorted_and_normalised(oracle, h2o_vdz):
    for i in range(h2o_vdz.nbas):
        e = h2o_vdz.bas_exp(i)
        assert numpy.all(numpy.diff(e) < 0) or len(e) == 1
    s = oracle.int1e(h2o_vdz, 'ovlp')
    assert abs(numpy.diag(s) - 1).max() < 1e-14
    assert abs(s - s.T).max() < 1e-15


def test_energy_nuc(pins, h2o_631g):
    assert abs(h2o_631g.energy_nuc() - pins['pin7_h2o_631g_energy_nuc']['value']) < 1e-10


def test_env_layout(h2o_vdz):
    atm = h2o_vdz._atm
    assert atm.shape == (3, 6) and list(atm[:, 0]) == [8, 1, 1]
    assert list(atm[:, 1]) == [20, 24, 28]
    assert abs(h2o_vdz._env[28:31] - numpy.array([0, 0.757, 0.587]) / 0.52917721092).max() < 1e-15
