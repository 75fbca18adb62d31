"""GPU J/K engine against the CPU oracle and the reference's golden values.

FP64 tolerance: 1e-10 max-abs on J and K (the target for the FP64 path)."""
import numpy
import pytest

from pyscf_amd import gto, lib, workloads
from pyscf_amd.scf import _vhf, hf
from tests.conftest import H2O

pytestmark = pytest.mark.gpu
TOL = 1e-10


def c1_density(nao=24):
    c = numpy.random.default_rng(0).standard_normal((nao, 5))
    return c @ c.T


def test_c1_h2o_ccpvdz_hermi1(oracle, h2o_vdz):
    dm = c1_density()
    vj, vk = hf.get_jk(h2o_vdz, dm, hermi=1)
    j0, k0 = oracle.get_jk(h2o_vdz, dm, hermi=1)
    assert vj.shape == dm.shape and vk.shape == dm.shape
    assert abs(vj - j0).max() < TOL
    assert abs(vk - k0).max() < TOL


def test_pin2_nonsymmetric_dm(oracle, pins, h2o_vdz):
    numpy.random.seed(1)
    dm = numpy.random.random((24, 24))
    vj, vk = hf.get_jk(h2o_vdz, dm, hermi=0)
    assert abs(numpy.linalg.norm(vj) - pins['pin2_h2o_ccpvdz_norm_J']['value']) < 1e-9
    assert abs(numpy.linalg.norm(vk) - pins['pin2_h2o_ccpvdz_norm_K']['value']) < 1e-9
    j0, k0 = oracle.get_jk(h2o_vdz, dm, hermi=0)
    assert abs(vj - j0).max() < TOL and abs(vk - k0).max() < TOL


def test_pin3_veff_two_densities(pins, h2o_vdz):
    numpy.random.seed(1)
    d1 = numpy.random.random((24, 24))
    d2 = numpy.random.random((24, 24))
    v = hf.get_veff(h2o_vdz, (d1 + d1.T, d2 + d2.T))
    assert v.shape == (2, 24, 24)
    assert abs(numpy.linalg.norm(v) - pins['pin3_h2o_ccpvdz_norm_veff']['value']) < 1e-9


def test_pin4_631g(pins, h2o_631g):
    numpy.random.seed(1)
    dm = numpy.random.random((13, 13))
    dm = dm + dm.T
    vj, vk = hf.get_jk(h2o_631g, dm, hermi=0)
    assert abs(lib.fp(vk) - pins['pin4_h2o_631g_fp_K_s1il']['value']) < 1e-11
    assert abs(lib.fp(vj) - pins['pin4_h2o_631g_fp_J_s1kl']['value']) < 1e-11


def test_pin5_complex_cart_screened(pins, oracle):
    mol = gto.M(atom=H2O, basis='631g', cart=True)
    opt = _vhf.VHFOpt(mol, direct_scf_tol=1e-13)
    numpy.random.seed(1)
    dm = numpy.random.random((2, 3, 13, 13)) - .5 + 0j
    vj, vk = hf.get_jk(mol, dm, hermi=0, vhfopt=opt)
    assert vj.shape == dm.shape and vj.dtype == numpy.complex128
    assert abs(lib.fp(vj - .5 * vk) - pins['pin5_h2o_631g_cart_fp_vhf4']['value']) < 1e-11
    vj2, vk2 = hf.get_jk(mol, dm[0, 0], hermi=0, vhfopt=opt)
    assert vj2.ndim == 2 and abs(vj2 - vj[0, 0]).max() < 1e-12


def test_flags_and_hermi2(oracle, h2o_vdz):
    rng = numpy.random.default_rng(3)
    a = rng.standard_normal((24, 24))
    vj, vk = hf.get_jk(h2o_vdz, a + a.T, with_k=False)
    assert vk is None
    j0, _ = oracle.get_jk(h2o_vdz, a + a.T)
    assert abs(vj - j0).max() < TOL
    vj, vk = hf.get_jk(h2o_vdz, a - a.T, hermi=2, with_j=False)
    assert vj is None
    _, k0 = oracle.get_jk(h2o_vdz, a - a.T, hermi=2)
    assert abs(vk - k0).max() < TOL


def test_schwarz_bound_matches_oracle(oracle, h2o_vdz):
    opt = _vhf.VHFOpt(h2o_vdz, direct_scf_tol=1e-13)
    q0 = oracle.VHFOpt(h2o_vdz, 1e-13).q_cond
    q = opt.q_cond
    assert abs(q - q0).max() < 1e-12 * abs(q0).max()


def test_screened_cluster(oracle):
    mol = gto.M(atom='''O 0 0 0; H 0 -0.757 0.587; H 0 0.757 0.587
                        O 3.1 0.2 0; H 3.1 -0.557 0.587; H 3.1 0.957 0.587
                        O 0 2.9 0.3; H 0 2.143 0.887; H 0 3.657 0.887''', basis='def2-svp')
    nao = mol.nao_nr()
    c = numpy.random.default_rng(5).standard_normal((nao, 15))
    dm = c @ c.T / 15
    opt = _vhf.VHFOpt(mol, direct_scf_tol=1e-13)
    vj, vk = hf.get_jk(mol, dm, vhfopt=opt)
    j0, k0 = oracle.get_jk(mol, dm, vhfopt=oracle.VHFOpt(mol, 1e-13))
    assert abs(vj - j0).max() < TOL and abs(vk - k0).max() < TOL
    j1, k1 = hf.get_jk(mol, dm)
    assert abs(vj - j1).max() < 1e-9 and abs(vk - k1).max() < 1e-9


def test_run_to_run_determinism(h2o_vdz):
    dm = c1_density()
    outs = [hf.get_jk(h2o_vdz, dm) for _ in range(3)]
    for vj, vk in outs[1:]:
        assert abs(vj - outs[0][0]).max() < 1e-13 and abs(vk - outs[0][1]).max() < 1e-13


def test_batch_equals_single(oracle):
    mols = [gto.M(atom=H2O, basis='cc-pvdz'), gto.M(atom='O 0 0 0; H 0 0 0.96; H 0.93 0 -0.25', basis='631g')]
    eng = _vhf.get_engine()
    basis = eng.upload_batch(mols, 0.0)
    dms = []
    for m in mols:
        c = numpy.random.default_rng(m.nao_nr()).standard_normal((m.nao_nr(), 4))
        dms.append(c @ c.T)
    flat = numpy.concatenate([d.ravel() for d in dms])[None]
    vj, vk = basis.get_jk(flat, 1, True, True, 0.0)
    off = 0
    for m, d in zip(mols, dms):
        n = m.nao_nr()
        j0, k0 = oracle.get_jk(m, d)
        assert abs(vj[0, off:off + n * n].reshape(n, n) - j0).max() < TOL
        assert abs(vk[0, off:off + n * n].reshape(n, n) - k0).max() < TOL
        off += n * n


def test_fshells_def2_tzvp(oracle):
    # O: 5s3p2d1f -- every class up to (ff|ff) via the cooperative kernel
    mol = gto.M(atom='O 0 0 0; H 0 -0.757 0.587; H 0 0.757 0.587; O 2.5 0.3 0.4', basis='def2-tzvp')
    nao = mol.nao_nr()
    rng = numpy.random.default_rng(7)
    c = rng.standard_normal((nao, 8))
    dm = c @ c.T / 8
    vj, vk = hf.get_jk(mol, dm, hermi=1)
    j0, k0 = oracle.get_jk(mol, dm, hermi=1)
    assert abs(vj - j0).max() < TOL and abs(vk - k0).max() < TOL
    a = rng.standard_normal((nao, nao))
    vj, vk = hf.get_jk(mol, a, hermi=0)
    j0, k0 = oracle.get_jk(mol, a, hermi=0)
    assert abs(vj - j0).max() < TOL * 10 and abs(vk - k0).max() < TOL * 10


def test_row_kernel_b_rows_in_memory(oracle):
    """With the row-kernel LDS limit just above the 10 A rows of an f-shell bra
    (and below its full 13-20 rows), the (fp|..), (fd|..) and (ff|ss) classes
    run the row kernel that keeps the B rows of K' in memory; J and K equal
    the oracle's, for one density (one K' set) and for a non-symmetric one
    (two)."""
    mol = gto.M(atom='O 0 0 0; H 0 -0.757 0.587; H 0 0.757 0.587; O 2.5 0.3 0.4', basis='def2-tzvp')
    nao = mol.nao_nr()
    bas = numpy.asarray(mol._bas)
    ncart = int(sum((l + 1) * (l + 2) // 2 * nc for l, nc in zip(bas[:, 1], bas[:, 3])))
    rng = numpy.random.default_rng(13)
    c = rng.standard_normal((nao, 8))
    eng = _vhf.get_engine()
    opt = _vhf.VHFOpt(mol, direct_scf_tol=1e-13)
    try:
        for hermi, n_k, dm in ((1, 1, c @ c.T / 8), (0, 2, rng.standard_normal((nao, nao)))):
            lds = 8 * (n_k * 10 * ncart + 10 * 10)
            assert eng.lib.mjk_ctx_set_row_lds(eng.handle, lds) == 0
            kinds = {cl[1]: cl[4] for cl in opt.basis.classes(1, hermi)}
            assert kinds[(3, 1, 0, 0)] == 4 and kinds[(3, 1, 1, 0)] == 4 and kinds[(3, 3, 0, 0)] == 4, kinds
            vj, vk = hf.get_jk(mol, dm, hermi=hermi, vhfopt=opt)
            j0, k0 = oracle.get_jk(mol, dm, hermi=hermi)
            print('B rows in memory, hermi=%d: max|dJ|=%.2e max|dK|=%.2e' % (hermi, abs(vj - j0).max(),
                                                                         abs(vk - k0).max()), flush=True)
            assert abs(vj - j0).max() < TOL * max(1, abs(j0).max()) and abs(vk - k0).max() < TOL * max(1, abs(k0).max())
    finally:
        assert eng.lib.mjk_ctx_set_row_lds(eng.handle, 162816) == 0


def test_cooperative_classes_dd(oracle):
    # (dd|dd) and (dd|dp) run the cooperative LDS kernel
    mol = gto.M(atom='O 0 0 0; O 0 0 1.2; O 1.1 0.2 0.3', basis='def2-svp')
    nao = mol.nao_nr()
    c = numpy.random.default_rng(11).standard_normal((nao, 12))
    dm = c @ c.T / 12
    opt = _vhf.VHFOpt(mol, direct_scf_tol=1e-13)
    vj, vk = hf.get_jk(mol, dm, vhfopt=opt)
    j0, k0 = oracle.get_jk(mol, dm, vhfopt=oracle.VHFOpt(mol, 1e-13))
    assert abs(vj - j0).max() < TOL and abs(vk - k0).max() < TOL


def test_cooperative_density_counts(oracle):
    """The cooperative kernel sums a bra pair's J'[a,b] over its kets in LDS for
    one or two densities and adds per quartet for more: J and K of each density
    agree whether it is built alone, in a pair or among three, and with the
    oracle; the row kernels fall back to the cooperative kernel (row LDS 0) so
    that most d-shell classes take it."""
    print('cooperative kernel, 1-3 densities', flush=True)
    mol = gto.M(atom='O 0 0 0; O 0 0 1.2; O 1.1 0.2 0.3', basis='def2-svp')
    nao = mol.nao_nr()
    rng = numpy.random.default_rng(12)
    dms = rng.standard_normal((3, nao, nao)) / 4
    opt = _vhf.VHFOpt(mol, direct_scf_tol=1e-13)
    eng = _vhf.get_engine()
    try:
        assert eng.lib.mjk_ctx_set_row_lds(eng.handle, 0) == 0
        single = [hf.get_jk(mol, dms[i], hermi=0, vhfopt=opt) for i in range(3)]
        pair = hf.get_jk(mol, dms[:2], hermi=0, vhfopt=opt)
        three = hf.get_jk(mol, dms, hermi=0, vhfopt=opt)
    finally:
        assert eng.lib.mjk_ctx_set_row_lds(eng.handle, 162816) == 0
    for i in range(3):
        js, ks = single[i]
        scale = max(1.0, abs(js).max(), abs(ks).max())
        assert abs(three[0][i] - js).max() < 1e-12 * scale and abs(three[1][i] - ks).max() < 1e-12 * scale
        if i < 2:
            assert abs(pair[0][i] - js).max() < 1e-12 * scale and abs(pair[1][i] - ks).max() < 1e-12 * scale
    j0, k0 = oracle.get_jk(mol, dms[2], hermi=0, vhfopt=oracle.VHFOpt(mol, 1e-13))
    assert abs(three[0][2] - j0).max() < TOL * max(1.0, abs(j0).max())
    assert abs(three[1][2] - k0).max() < TOL * max(1.0, abs(k0).max())


@pytest.mark.parametrize('row_lds', [0, 3072])
def test_row_and_tile_kernels_agree(oracle, row_lds):
    # row kernels (K' rows in LDS) against their fallback kernels: 0 forces the fallback (tile or
    # cooperative) everywhere, 3072 bytes keeps row kernels only for the narrowest classes
    mol = gto.M(atom='''O 0 0 0; H 0 -0.757 0.587; H 0 0.757 0.587
                        O 3.1 0.2 0; H 3.1 -0.557 0.587; H 3.1 0.957 0.587''', basis='def2-svp')
    nao = mol.nao_nr()
    rng = numpy.random.default_rng(13)
    c = rng.standard_normal((nao, 10))
    dms = numpy.array([c @ c.T / 10, rng.standard_normal((nao, nao))])
    opt = _vhf.VHFOpt(mol, direct_scf_tol=1e-13)
    eng = _vhf.get_engine()
    ref = [hf.get_jk(mol, dms[0], hermi=1, vhfopt=opt), hf.get_jk(mol, dms, hermi=0, vhfopt=opt)]
    assert eng.lib.mjk_ctx_set_row_lds(eng.handle, row_lds) == 0
    try:
        got = [hf.get_jk(mol, dms[0], hermi=1, vhfopt=opt), hf.get_jk(mol, dms, hermi=0, vhfopt=opt)]
    finally:
        assert eng.lib.mjk_ctx_set_row_lds(eng.handle, 162816) == 0
    for (j0, k0), (j1, k1) in zip(ref, got):
        assert abs(j1 - j0).max() < 1e-11 * max(1, abs(j0).max())
        assert abs(k1 - k0).max() < 1e-11 * max(1, abs(k0).max())
    j0, k0 = oracle.get_jk(mol, dms, hermi=0, vhfopt=oracle.VHFOpt(mol, 1e-13))
    assert abs(got[1][0] - j0).max() < TOL * 10 and abs(got[1][1] - k0).max() < TOL * 10


def test_quartet_count_matches_oracle(oracle):
    # segmented basis (no general contractions): the engine evaluates exactly the
    # shell quartets the reference's prescreen keeps, each counted once however
    # its primitive pairs are shared out among row tasks
    mol = gto.M(atom='''O 0 0 0; H 0 -0.757 0.587; H 0 0.757 0.587
                        O 3.1 0.2 0; H 3.1 -0.557 0.587; H 3.1 0.957 0.587
                        O 0 2.9 0.3; H 0 2.143 0.887; H 0 3.657 0.887''', basis='def2-svp')
    nao = mol.nao_nr()
    c = numpy.random.default_rng(21).standard_normal((nao, 15))
    dm = c @ c.T / 15
    opt = _vhf.VHFOpt(mol, direct_scf_tol=1e-13)
    opt.basis.get_jk(dm[None], 1, True, True, 1e-13, stats=True)
    nq = opt.basis.last_stats['quartets']
    _, _, nq0 = oracle.direct_full(mol, dm[None], vhfopt=oracle.VHFOpt(mol, 1e-13), pair_stride=0)
    assert abs(nq - nq0) <= max(2, 1e-4 * nq0), (nq, nq0)


@pytest.mark.parametrize('which', ['j', 'k'])
def test_j_only_k_only_prescreens(oracle, which):
    """J-only and K-only builds screen with CVHFnrs8_vj_prescreen /
    CVHFnrs8_vk_prescreen: the engine evaluates the quartets those predicates
    keep, and the matrix agrees with the unscreened one."""
    print('prescreen', which, flush=True)
    mol = gto.M(atom='''O 0 0 0; H 0 -0.757 0.587; H 0 0.757 0.587
                        O 3.1 0.2 0; H 3.1 -0.557 0.587; H 3.1 0.957 0.587
                        O 0 2.9 0.3; H 0 2.143 0.887; H 0 3.657 0.887''', basis='def2-svp')
    nao = mol.nao_nr()
    c = numpy.random.default_rng(23).standard_normal((nao, 15))
    dm = c @ c.T / 15
    opt = _vhf.VHFOpt(mol, direct_scf_tol=1e-11)
    wj, wk = which == 'j', which == 'k'
    vj, vk = opt.basis.get_jk(dm[None], 1, wj, wk, 1e-11, stats=True)
    nq = opt.basis.last_stats['quartets']
    opt.basis.get_jk(dm[None], 1, True, True, 1e-11, stats=True)
    nq_both = opt.basis.last_stats['quartets']
    name = 'CVHFnrs8_vj_prescreen' if wj else 'CVHFnrs8_vk_prescreen'
    _, _, nq0 = oracle.direct_full(mol, dm[None], vhfopt=oracle.VHFOpt(mol, 1e-11, prescreen=name),
                                   pair_stride=0)
    assert abs(nq - nq0) <= max(2, 1e-4 * nq0), (nq, nq0)
    assert nq < nq_both
    j0, k0 = oracle.get_jk(mol, dm, hermi=1)
    if wj:
        assert vk is None and abs(vj[0] - j0).max() < 1e-8
    else:
        assert vj is None and abs(vk[0].reshape(nao, nao) - k0).max() < 1e-8


def test_graph_replays_match_eager(oracle):
    """Plain builds replay a captured HIP graph; builds with statistics run
    eagerly.  Several replays, each with eager builds of another density and
    synchronisations between them, every replay checked against the eager
    result of the same density."""
    print('graph replays', flush=True)
    mol = gto.M(atom='''O 0 0 0; H 0 -0.757 0.587; H 0 0.757 0.587
                        O 3.1 0.2 0; H 3.1 -0.557 0.587; H 3.1 0.957 0.587''', basis='def2-svp')
    nao = mol.nao_nr()
    rng = numpy.random.default_rng(29)
    c1, c2 = rng.standard_normal((nao, 10)), rng.standard_normal((nao, 10))
    dms = [c1 @ c1.T / 10, c2 @ c2.T / 10]
    opt = _vhf.VHFOpt(mol, direct_scf_tol=1e-13)
    eng = opt.engine
    ref = [opt.basis.get_jk(d[None], 1, True, True, 1e-13, stats=True) for d in dms]
    assert eng.lib.mjk_ctx_set_graphs(eng.handle, 1) == 0
    try:
        _graph_replays(oracle, mol, opt, dms, ref)
    finally:
        assert eng.lib.mjk_ctx_set_graphs(eng.handle, 0) == 0


def _graph_replays(oracle, mol, opt, dms, ref):
    for rep in range(4):
        for i in (0, 1):
            j, k = opt.basis.get_jk(dms[i][None], 1, True, True, 1e-13)          # graph
            assert abs(j - ref[i][0]).max() < 1e-12 * abs(ref[i][0]).max(), (rep, i)
            assert abs(k - ref[i][1]).max() < 1e-12 * abs(ref[i][1]).max(), (rep, i)
            je, ke = opt.basis.get_jk(dms[1 - i][None], 1, True, True, 1e-13, stats=True)   # eager between
            assert abs(je - ref[1 - i][0]).max() < 1e-12 * abs(ref[1 - i][0]).max()
    # a different request (J only) captures its own graph; the old one is replaced
    j, k = opt.basis.get_jk(dms[0][None], 1, True, False, 1e-13)
    assert k is None and abs(j - ref[0][0]).max() < 1e-9
    j, k = opt.basis.get_jk(dms[0][None], 1, True, True, 1e-13)
    assert abs(k - ref[0][1]).max() < 1e-12 * abs(ref[0][1]).max()
    j0, k0 = oracle.get_jk(mol, dms[0], hermi=1)
    assert abs(j - j0).max() < 1e-10 and abs(k - k0).max() < 1e-10


def test_row_kernels_above_64k_lds(oracle):
    """Four non-symmetric densities give the (dd|..) row kernels K' rows of
    8 x 12 x ld doubles, more than 64 KB: they run with the 160 KB limit and
    agree with the fallback kernels and the oracle."""
    print('row kernels > 64 KB LDS', flush=True)
    mol = gto.M(atom=workloads.water_cluster(4, seed=3), basis='def2-svp')
    nao = mol.nao_nr()
    rng = numpy.random.default_rng(31)
    dms = rng.standard_normal((4, nao, nao))
    opt = _vhf.VHFOpt(mol, direct_scf_tol=1e-13)
    eng = _vhf.get_engine()
    out = {}
    try:
        for lds in (162816, 0):
            assert eng.lib.mjk_ctx_set_row_lds(eng.handle, lds) == 0
            out[lds] = hf.get_jk(mol, dms, hermi=0, vhfopt=opt)
    finally:
        assert eng.lib.mjk_ctx_set_row_lds(eng.handle, 162816) == 0
    (j1, k1), (j0, k0) = out[162816], out[0]
    assert abs(j1 - j0).max() < 1e-11 * max(1, abs(j0).max())
    assert abs(k1 - k0).max() < 1e-11 * max(1, abs(k0).max())
    jo, ko = oracle.get_jk(mol, dms[1], hermi=0)
    assert abs(j1[1] - jo).max() < 1e-9 and abs(k1[1] - ko).max() < 1e-9
