"""Host one-electron integrals of the library against the oracle."""
import numpy
import pytest

from pyscf_amd import gto
from pyscf_amd.scf import rhf
from tests.conftest import H2O


@pytest.mark.parametrize('basis', ['631g', 'cc-pvdz', 'def2-tzvp'])
def test_int1e_matches_oracle(oracle, basis):
    mol = gto.M(atom=H2O, basis=basis)
    for which in ('ovlp', 'kin', 'nuc'):
        a = rhf.int1e(mol, which)
        b = oracle.int1e(mol, which)
        assert abs(a - b).max() < 1e-12, which
    s = rhf.int1e(mol, 'ovlp')
    assert abs(numpy.diag(s) - 1).max() < 1e-13
