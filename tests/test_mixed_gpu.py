"""Mixed precision: FP32 recurrences only for waves of small quartets; J/K
error and converged RHF energy stay within the stated tolerances."""
import numpy
import pytest

from pyscf_amd import gto, workloads
from pyscf_amd.scf import _vhf, hf, rhf

pytestmark = pytest.mark.gpu
FP32_TOL = 1e-7


def cluster(n):
    return gto.M(atom=workloads.water_cluster(n, seed=1), basis='def2-svp')


def test_mixed_jk_error_bounded():
    mol = cluster(4)
    nao = mol.nao_nr()
    c = numpy.random.default_rng(2).standard_normal((nao, 20)) / 5
    dm = c @ c.T
    opt = _vhf.VHFOpt(mol, direct_scf_tol=1e-13)
    j64, k64 = hf.get_jk(mol, dm, vhfopt=opt)
    j32, k32 = hf.get_jk(mol, dm, vhfopt=opt, fp32_tol=FP32_TOL)
    st = opt.basis
    assert abs(j32 - j64).max() < 1e-8
    assert abs(k32 - k64).max() < 1e-8
    # the FP32 path was actually taken for part of the work
    opt.basis.get_jk(dm[None], 1, True, True, 1e-13, FP32_TOL, stats=True)
    assert opt.basis.last_stats['quartets_fp32'] > 0


def test_mixed_rhf_energy():
    mol = cluster(4)
    e = []
    for tol in (0.0, FP32_TOL):
        mf = rhf.RHF(mol, fp32_tol=tol)
        mf.conv_tol = 1e-10
        e.append(mf.kernel())
        assert mf.converged
    assert abs(e[1] - e[0]) < 1e-8


def test_fp32_tol_zero_is_fp64(h2o_vdz):
    dm = numpy.eye(24) * 0.1
    opt = _vhf.VHFOpt(h2o_vdz, direct_scf_tol=1e-13)
    a = hf.get_jk(h2o_vdz, dm, vhfopt=opt)
    b = hf.get_jk(h2o_vdz, dm, vhfopt=opt, fp32_tol=0.0)
    assert abs(a[0] - b[0]).max() == 0 or abs(a[0] - b[0]).max() < 1e-14
