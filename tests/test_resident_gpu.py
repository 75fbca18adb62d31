"""Resident builds (mjk_get_jk_resident / mjk_basis_fetch_jk) and the
quartet-class filter (mjk_basis_set_class_filter), through the C ABI.

* A resident build of the density the last get_jk uploaded equals that
  get_jk's J and K (the same kernels; summation order of the atomics aside).
* J and K are linear in the quartets: the builds restricted to each class
  add up to the full build, and their evaluated-quartet counts add up to the
  full count exactly (FP64 and mixed precision).
* Refusals: a resident build with no resident density, a class index out of
  range.
"""
import numpy
import pytest

from pyscf_amd import gto, workloads
from pyscf_amd.scf import _vhf

pytestmark = pytest.mark.gpu
TOL = 1e-10
SCF_TOL = 1e-13


def density(nao, seed):
    c = numpy.random.default_rng(seed).standard_normal((nao, max(2, nao // 4))) / 4
    return c @ c.T


@pytest.fixture(scope='module')
def water4():
    mol = gto.M(atom=workloads.water_cluster(4, seed=3), basis='def2-svp')
    opt = _vhf.VHFOpt(mol, direct_scf_tol=SCF_TOL)
    return mol, opt


def test_resident_equals_host_build(water4):
    mol, opt = water4
    b = opt.basis
    dm = density(mol.nao_nr(), 21)
    vj, vk = (a.reshape(1, -1) for a in b.get_jk(dm[None], 1, True, True, SCF_TOL))
    for _ in range(3):               # eager, captured, replayed
        b.get_jk_resident(1, 1, True, True, SCF_TOL)
        rj, rk = b.fetch_jk(1)
        assert abs(rj - vj).max() <= 1e-12 and abs(rk - vk).max() <= 1e-12
    assert abs(vj).max() > 1e-2 and abs(vk).max() > 1e-2


@pytest.mark.parametrize('fp32_tol', [0.0, 1e-7])
def test_class_builds_add_up(water4, fp32_tol):
    mol, opt = water4
    b = opt.basis
    dm = density(mol.nao_nr(), 22)
    vj, vk = (a.reshape(1, -1) for a in b.get_jk(dm[None], 1, True, True, SCF_TOL, fp32_tol))
    full = b.get_jk_resident(1, 1, True, True, SCF_TOL, fp32_tol, stats=True)
    sj, sk = numpy.zeros_like(vj), numpy.zeros_like(vk)
    nq = nq32 = 0
    classes = [c for c in b.classes() if c[3] > 0 and c[4] >= 0]
    assert len(classes) >= 10
    try:
        for k, l4, ctr, nblocks, kernel in classes:
            assert kernel in (0, 1, 2, 3)
            b.set_class_filter(k)
            st = b.get_jk_resident(1, 1, True, True, SCF_TOL, fp32_tol, stats=True)
            cj, ck = b.fetch_jk(1)
            sj += cj
            sk += ck
            nq += st['quartets']
            nq32 += st['quartets_fp32']
            assert st['ms_quartets'] > 0
    finally:
        b.set_class_filter(-1)
    assert nq == full['quartets'] and nq32 == full['quartets_fp32']
    assert abs(sj - vj).max() <= TOL and abs(sk - vk).max() <= TOL
    # the filter is off again: a plain build is the full build
    b.get_jk_resident(1, 1, True, True, SCF_TOL, fp32_tol)
    rj, rk = b.fetch_jk(1)
    assert abs(rj - vj).max() <= 1e-12 and abs(rk - vk).max() <= 1e-12


def test_refusals():
    mol = gto.M(atom=workloads.h2o(), basis='cc-pvdz')
    opt = _vhf.VHFOpt(mol, direct_scf_tol=SCF_TOL)
    b = opt.basis
    with pytest.raises(RuntimeError, match='no resident density'):
        b.get_jk_resident(1, 1, True, True, SCF_TOL)
    with pytest.raises(RuntimeError, match='out of range'):
        b.set_class_filter(len(b.classes()))
    with pytest.raises(RuntimeError, match='out of range'):
        b.set_class_filter(-2)
