"""Generally contracted s shells kept whole (merged pairs).

cc-pVDZ C, N and O carry an s shell of two contractions over the same eight
primitives.  The engine evaluates such shells' quartets once for both
contractions (merged row kernels) when the class's merged kernels exist and
their LDS fits, else through the split pairs on the plain kernels.  Checks:

* J and K against the oracle at the FP64 contract (1e-10 max-abs), for one
  symmetric density and for two non-symmetric densities (hermi=0, twice the
  K' rows in LDS);
* some classes really run on the merged kernel, and the same build with the
  merged kernels' LDS taken away (row LDS 0: split pairs, fallback kernels)
  agrees with it;
* mixed precision on the merged kernels stays within 1e-8 of FP64.
"""
import numpy
import pytest

from pyscf_amd import gto, workloads
from pyscf_amd.scf import _vhf

pytestmark = pytest.mark.gpu
TOL = 1e-10
SCF_TOL = 1e-13

MOLECULES = {
    'water3': lambda: gto.M(atom=workloads.water_cluster(3, seed=4), basis='cc-pvdz'),
    'butane': lambda: gto.M(atom=workloads.alkane(4), basis='cc-pvdz'),
    'chain': lambda: gto.M(atom=workloads.random_chain([7, 1], nmin=6, nmax=6), basis='cc-pvdz'),
}


def densities(nao, n_dm, seed, symmetric):
    rng = numpy.random.default_rng(seed)
    out = []
    for _ in range(n_dm):
        c = rng.standard_normal((nao, max(2, nao // 4))) / 4
        d = c @ c.T
        if not symmetric:
            d = d + 0.05 * rng.standard_normal((nao, nao))
        out.append(d)
    return numpy.array(out)


@pytest.fixture(scope='module')
def fast_oracle(oracle):
    oracle.set_roots(1)
    yield oracle
    oracle.set_roots(0)


@pytest.mark.parametrize('name', sorted(MOLECULES))
def test_merged_vs_oracle(fast_oracle, name):
    mol = MOLECULES[name]()
    nao = mol.nao_nr()
    print('merged', name, nao, flush=True)
    opt = _vhf.VHFOpt(mol, direct_scf_tol=SCF_TOL)
    b = opt.basis
    kinds = [c[4] for c in b.classes() if c[3] > 0]
    assert 3 in kinds, 'no class runs on the merged kernel'
    orc = fast_oracle.VHFOpt(mol, SCF_TOL)
    for n_dm, hermi, seed in ((1, 1, 31), (2, 0, 32)):
        dms = densities(nao, n_dm, seed, hermi == 1)
        vj, vk = b.get_jk(dms, hermi, True, True, SCF_TOL)
        j0, k0, _ = fast_oracle.direct_full(mol, dms, vhfopt=orc)
        if hermi == 1:
            k0 = fast_oracle.hermi_triu(k0.copy(), 1)
        dj, dk = abs(vj - j0).max(), abs(vk - k0).max()
        print('  n_dm=%d hermi=%d max|dJ|=%.2e max|dK|=%.2e' % (n_dm, hermi, dj, dk), flush=True)
        assert dj <= TOL and dk <= TOL


def test_merged_vs_split_and_mixed():
    mol = MOLECULES['butane']()
    nao = mol.nao_nr()
    eng = _vhf.get_engine()
    opt = _vhf.VHFOpt(mol, direct_scf_tol=SCF_TOL)
    b = opt.basis
    dm = densities(nao, 1, 33, True)
    vj, vk = b.get_jk(dm, 1, True, True, SCF_TOL)
    lds = 162816
    try:
        eng.check(eng.lib.mjk_ctx_set_row_lds(eng.handle, 0), 'mjk_ctx_set_row_lds')
        assert 3 not in [c[4] for c in b.classes() if c[3] > 0]
        sj, sk = b.get_jk(dm, 1, True, True, SCF_TOL)
    finally:
        eng.check(eng.lib.mjk_ctx_set_row_lds(eng.handle, lds), 'mjk_ctx_set_row_lds')
    print('merged vs split: max|dJ|=%.2e max|dK|=%.2e' % (abs(vj - sj).max(), abs(vk - sk).max()), flush=True)
    assert abs(vj - sj).max() <= TOL and abs(vk - sk).max() <= TOL
    mj, mk = b.get_jk(dm, 1, True, True, SCF_TOL, fp32_tol=1e-7)
    assert abs(mj - vj).max() <= 1e-8 and abs(mk - vk).max() <= 1e-8
