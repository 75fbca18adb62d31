"""Edge cases of the GPU J/K build against the CPU oracle: one atom, s-only
bases, zero and several densities, far-apart fragments (most quartets
screened), a range-separated operator on a single atom.

FP64 tolerance: 1e-10 max-abs on J and K."""
import numpy
import pytest

from pyscf_amd import gto
from pyscf_amd.scf import _vhf, hf

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _density(nao, nocc, seed):
    c = numpy.random.default_rng(seed).standard_normal((nao, nocc))
    return c @ c.T


@pytest.mark.parametrize('omega', [None, -0.5])
def test_single_atom(oracle, omega):
    print('edge single atom omega', omega, flush=True)
    mol = gto.M(atom='He 0 0 0', basis='cc-pvdz')
    dm = _density(mol.nao_nr(), 1, 1)
    vj, vk = hf.get_jk(mol, dm, hermi=1, omega=omega)
    with mol.with_range_coulomb(omega):
        j0, k0 = oracle.get_jk(mol, dm, hermi=1)
    assert abs(vj - j0).max() < TOL and abs(vk - k0).max() < TOL


def test_s_only_basis(oracle):
    """6-31G hydrogen chain: only s shells, so every class but (ss|ss) is empty."""
    print('edge s-only', flush=True)
    mol = gto.M(atom='H 0 0 0; H 0 0 0.74; H 0 0 2.5; H 0 0 3.24', basis='631g')
    dm = _density(mol.nao_nr(), 2, 2)
    opt = _vhf.VHFOpt(mol, direct_scf_tol=1e-13)
    vj, vk = hf.get_jk(mol, dm, hermi=1, vhfopt=opt)
    j0, k0 = oracle.get_jk(mol, dm, hermi=1)
    assert abs(vj - j0).max() < TOL and abs(vk - k0).max() < TOL


def test_zero_density(h2o_vdz):
    print('edge zero density', flush=True)
    dm = numpy.zeros((24, 24))
    opt = _vhf.VHFOpt(h2o_vdz, direct_scf_tol=1e-13)
    for o in (None, opt):
        vj, vk = hf.get_jk(h2o_vdz, dm, hermi=1, vhfopt=o)
        assert not vj.any() and not vk.any()


def test_three_densities(oracle, h2o_vdz):
    print('edge three densities', flush=True)
    dms = numpy.array([_density(24, 5, s) for s in (3, 4, 5)])
    vj, vk = hf.get_jk(h2o_vdz, dms, hermi=1)
    assert vj.shape == (3, 24, 24)
    for i in range(3):
        j0, k0 = oracle.get_jk(h2o_vdz, dms[i], hermi=1)
        assert abs(vj[i] - j0).max() < TOL and abs(vk[i] - k0).max() < TOL


def test_far_apart_fragments(oracle):
    """Two waters 30 A apart: the Schwarz prefix drops every inter-fragment
    exchange-type quartet; the screened build still matches the unscreened
    oracle."""
    print('edge far fragments', flush=True)
    mol = gto.M(atom='''O 0 0 0; H 0 -0.757 0.587; H 0 0.757 0.587
                        O 30 0 0; H 30 -0.757 0.587; H 30 0.757 0.587''', basis='def2-svp')
    nao = mol.nao_nr()
    dm = _density(nao, 10, 6)
    opt = _vhf.VHFOpt(mol, direct_scf_tol=1e-13)
    opt.basis.get_jk(dm[None], 1, True, True, 1e-13, stats=True)
    nq = opt.basis.last_stats['quartets']
    vj, vk = hf.get_jk(mol, dm, hermi=1, vhfopt=opt)
    j0, k0 = oracle.get_jk(mol, dm, hermi=1)
    assert abs(vj - j0).max() < 1e-9 and abs(vk - k0).max() < 1e-9
    nbas = mol.nbas
    npair = nbas * (nbas + 1) // 2
    assert nq < 0.5 * npair * (npair + 1) // 2    # most quartets screened
