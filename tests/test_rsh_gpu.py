"""Range-separated Coulomb J/K on the GPU against the CPU oracle.

The operator comes from env[PTR_RANGE_OMEGA] of the uploaded basis (omega > 0
erf(omega r)/r, omega < 0 erfc(|omega| r)/r), as ``mol.with_range_coulomb``
sets it (pyscf/scf/hf.py:1006).  FP64 tolerance: 1e-10 max-abs on J and K."""
import numpy
import pytest

from pyscf_amd import gto
from pyscf_amd.scf import _vhf, hf, rhf
from tests.conftest import H2O

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.mark.parametrize('omega', [0.4, -0.4])
def test_c1_erf_erfc_nonsymmetric(oracle, h2o_vdz, omega):
    print('rsh c1 omega', omega, flush=True)
    rng = numpy.random.default_rng(11)
    dm = rng.standard_normal((24, 24))
    vj, vk = hf.get_jk(h2o_vdz, dm, hermi=0, omega=omega)
    assert h2o_vdz.omega == 0.0
    with h2o_vdz.with_range_coulomb(omega):
        j0, k0 = oracle.get_jk(h2o_vdz, dm, hermi=0)
    assert abs(vj - j0).max() < TOL
    assert abs(vk - k0).max() < TOL


def test_c1_erf_plus_erfc_is_coulomb(h2o_vdz):
    print('rsh c1 erf + erfc', flush=True)
    c = numpy.random.default_rng(0).standard_normal((24, 5))
    dm = c @ c.T
    j0, k0 = hf.get_jk(h2o_vdz, dm)
    j1, k1 = hf.get_jk(h2o_vdz, dm, omega=0.25)
    j2, k2 = hf.get_jk(h2o_vdz, dm, omega=-0.25)
    assert abs(j1 + j2 - j0).max() < TOL and abs(k1 + k2 - k0).max() < TOL
    assert abs(k1 - k0).max() > 1e-2


@pytest.mark.parametrize('omega', [0.33, -0.33])
def test_tzvp_f_shells(oracle, omega):
    """def2-TZVP water: f shells, so the cooperative kernels take part."""
    print('rsh tzvp omega', omega, flush=True)
    mol = gto.M(atom=H2O, basis='def2-tzvp')
    nao = mol.nao_nr()
    c = numpy.random.default_rng(5).standard_normal((nao, 6))
    dm = c @ c.T
    vj, vk = hf.get_jk(mol, dm, hermi=1, omega=omega)
    with mol.with_range_coulomb(omega):
        j0, k0 = oracle.get_jk(mol, dm, hermi=1)
    assert abs(vj - j0).max() < TOL
    assert abs(vk - k0).max() < TOL


def test_screened_per_omega_optimizer(oracle, h2o_vdz):
    """RHF.get_jk(omega=...) builds its bounds under that operator: q_cond
    matches the oracle's erf bounds, and screened J/K match unscreened ones."""
    print('rsh screened optimizer', flush=True)
    mf = rhf.RHF(h2o_vdz, direct_scf_tol=1e-13)
    c = numpy.random.default_rng(2).standard_normal((24, 5))
    dm = c @ c.T
    vj, vk = mf.get_jk(h2o_vdz, dm, omega=0.3)
    opt = mf._opt[0.3]
    assert opt.omega == 0.3
    with h2o_vdz.with_range_coulomb(0.3):
        q0 = oracle.VHFOpt(h2o_vdz, 1e-13).q_cond
        j0, k0 = oracle.get_jk(h2o_vdz, dm, hermi=1)
    assert abs(opt.q_cond - q0).max() < 1e-12
    q_full = oracle.VHFOpt(h2o_vdz, 1e-13).q_cond
    assert numpy.all(q0 <= q_full + 1e-14)
    assert abs(vj - j0).max() < 1e-9 and abs(vk - k0).max() < 1e-9
    # the Coulomb optimizer is a different one and still gives Coulomb J/K
    vj0, _ = mf.get_jk(h2o_vdz, dm)
    assert abs(vj0 - vj).max() > 1e-2
    with pytest.raises(ValueError):
        hf.get_jk(h2o_vdz, dm, vhfopt=opt)          # bounds of erf used for 1/r


@pytest.mark.parametrize('omega', [0.3, -0.3])
def test_row_and_fallback_kernels_agree(oracle, omega):
    """Range-separated row kernels against their fallback (tile and
    cooperative) kernels, and both against the oracle."""
    print('rsh row vs fallback omega', omega, flush=True)
    mol = gto.M(atom='''O 0 0 0; H 0 -0.757 0.587; H 0 0.757 0.587
                        O 3.1 0.2 0; H 3.1 -0.557 0.587; H 3.1 0.957 0.587''', basis='def2-svp')
    nao = mol.nao_nr()
    rng = numpy.random.default_rng(17)
    dm = rng.standard_normal((nao, nao))
    with mol.with_range_coulomb(omega):
        opt = _vhf.VHFOpt(mol, direct_scf_tol=1e-13)
    eng = _vhf.get_engine()
    j_row, k_row = hf.get_jk(mol, dm, hermi=0, vhfopt=opt, omega=omega)
    assert eng.lib.mjk_ctx_set_row_lds(eng.handle, 0) == 0
    try:
        j_fb, k_fb = hf.get_jk(mol, dm, hermi=0, vhfopt=opt, omega=omega)
    finally:
        assert eng.lib.mjk_ctx_set_row_lds(eng.handle, 98304) == 0
    assert abs(j_row - j_fb).max() < 1e-11 * max(1, abs(j_fb).max())
    assert abs(k_row - k_fb).max() < 1e-11 * max(1, abs(k_fb).max())
    with mol.with_range_coulomb(omega):
        j0, k0 = oracle.get_jk(mol, dm, hermi=0, vhfopt=oracle.VHFOpt(mol, 1e-13))
    assert abs(j_row - j0).max() < TOL * 10 and abs(k_row - k0).max() < TOL * 10


def test_mixed_precision_and_batches(oracle):
    """The range-separated operators with FP32 recurrences for small quartets,
    in a batch of two molecules; a batch whose molecules disagree on omega is
    refused."""
    print('rsh mixed + batch', flush=True)
    m1 = gto.M(atom=H2O, basis='def2-svp')
    m2 = gto.M(atom='O 0 0 0; H 0 0 0.96; H 0.93 0 -0.25', basis='631g')
    eng = _vhf.get_engine()
    for m in (m1, m2):
        m.omega = 0.3
    b = eng.upload_batch([m1, m2], 1e-13)
    dms = [_dm(m, 5, 3 + i) for i, m in enumerate((m1, m2))]
    flat = numpy.zeros((1, b.slot))      # the molecules' nao^2 blocks back to back
    flat[0, :dms[0].size] = dms[0].ravel()
    flat[0, dms[0].size:dms[0].size + dms[1].size] = dms[1].ravel()
    vj, vk = b.get_jk(flat, 1, True, True, 1e-13, fp32_tol=1e-7)
    off = 0
    for m, d in zip((m1, m2), dms):
        n = d.size
        j0, k0 = oracle.get_jk(m, d, hermi=1)
        assert abs(vj[0, off:off + n].reshape(d.shape) - j0).max() < 1e-8
        assert abs(vk[0, off:off + n].reshape(d.shape) - k0).max() < 1e-8
        off += n
    m2.omega = -0.3
    with pytest.raises(RuntimeError):
        eng.upload_batch([m1, m2], 1e-13)
    for m in (m1, m2):
        m.omega = 0.0


def _dm(mol, nocc, seed):
    c = numpy.random.default_rng(seed).standard_normal((mol.nao_nr(), nocc))
    return c @ c.T
