"""The CPU oracle against the reference's golden values and closed forms."""
import ctypes
import math

import numpy
import pytest

from pyscf_amd import gto
from tests.conftest import H2O

TS = [0.0, 1e-12, 1e-6, 0.05, 0.5, 1.0, 1.999999, 2.0, 3.7, 9.99, 17.2, 29.5, 41.9, 42.0, 47.3,
      53.9, 60.1, 77.9, 78.0, 89.99, 90.0, 150.0, 1e3, 1e6]


@pytest.mark.parametrize('n', range(1, 10))
def test_rys_roots_reproduce_boys_moments(oracle, n):
    for T in TS:
        x, w = oracle.rys_roots(n, T)
        assert numpy.all(numpy.diff(x) > 0) and numpy.all((x > 0) & (x < 1)) and numpy.all(w > 0)
        for m in range(2 * n):
            ref = float(oracle.boys_mp(m, T))
            assert abs(numpy.dot(w, x**m) - ref) <= 1e-13 * ref, (n, T, m)


def test_boys_closed_forms(oracle):
    # F_0(T) = sqrt(pi/T)/2 erf(sqrt(T)), F_m(0) = 1/(2m+1)
    for T in (0.3, 2.0, 25.0):
        assert abs(float(oracle.boys_mp(0, T)) - 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))) < 1e-15
    for m in range(8):
        assert float(oracle.boys_mp(m, 0)) == pytest.approx(1.0 / (2 * m + 1), rel=1e-15)


def test_ssss_closed_form(oracle):
    # two uncontracted s functions per atom pair: (ss|ss) from the textbook formula
    a, b = 1.3, 0.45
    mol = gto.M(atom='H 0 0 0; H 0.3 -0.2 1.1', unit='Bohr',
                basis={'H': [[0, [a, 1.0]], [0, [b, 1.0]]]})
    eri = oracle.int2e_full(mol)
    A = numpy.zeros(3)
    B = numpy.array([0.3, -0.2, 1.1])
    exps = [(a, A), (b, A), (a, B), (b, B)]
    norm = [(2 * e / math.pi) ** 0.75 for e, _ in exps]

    def ssss(i, j, k, l):
        (ea, ra), (eb, rb), (ec, rc), (ed, rd) = exps[i], exps[j], exps[k], exps[l]
        p, q = ea + eb, ec + ed
        P = (ea * ra + eb * rb) / p
        Q = (ec * rc + ed * rd) / q
        kab = math.exp(-ea * eb / p * numpy.dot(ra - rb, ra - rb))
        kcd = math.exp(-ec * ed / q * numpy.dot(rc - rd, rc - rd))
        T = p * q / (p + q) * numpy.dot(P - Q, P - Q)
        f0 = 1.0 if T < 1e-15 else 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
        return 2 * math.pi**2.5 / (p * q * math.sqrt(p + q)) * kab * kcd * f0 * norm[i] * norm[j] * norm[k] * norm[l]

    for idx in [(0, 0, 0, 0), (0, 1, 2, 3), (1, 3, 0, 2), (3, 3, 1, 1)]:
        assert eri[idx] == pytest.approx(ssss(*idx), rel=1e-13)


def test_eri_permutational_symmetry(oracle, h2o_631g):
    eri = oracle.int2e_full(h2o_631g)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert abs(eri - eri.transpose(perm)).max() < 1e-13


def test_pin1_int2e_s8_fingerprint(oracle, pins):
    p = pins['pin1_hene_ccpvdz_int2e_s8_fp']
    mol = gto.M(atom=p['mol'], basis=p['basis'])
    eri = oracle.int2e_full(mol)
    assert abs(oracle.fp(oracle.pack_s8(eri)) - p['value']) < 0.5 * 10**-p['places']


def test_pin2_jk_norms_nonsymmetric_dm(oracle, pins, h2o_vdz):
    numpy.random.seed(1)
    dm = numpy.random.random((24, 24))
    vj, vk = oracle.get_jk(h2o_vdz, dm, hermi=0)
    assert abs(numpy.linalg.norm(vj) - pins['pin2_h2o_ccpvdz_norm_J']['value']) < 1e-9
    assert abs(numpy.linalg.norm(vk) - pins['pin2_h2o_ccpvdz_norm_K']['value']) < 1e-9
    eri = oracle.int2e_full(h2o_vdz)
    j0, k0 = oracle.jk_einsum(eri, dm)
    assert abs(vj - j0).max() < 1e-12 and abs(vk - k0).max() < 1e-12


def test_pin3_veff_norm(oracle, pins, h2o_vdz):
    numpy.random.seed(1)
    d1 = numpy.random.random((24, 24))
    d2 = numpy.random.random((24, 24))
    d = numpy.array((d1 + d1.T, d2 + d2.T))
    vj, vk = oracle.get_jk(h2o_vdz, d, hermi=1)
    assert abs(numpy.linalg.norm(vj - .5 * vk) - pins['pin3_h2o_ccpvdz_norm_veff']['value']) < 1e-9


def test_pin4_631g_fingerprints(oracle, pins, h2o_631g):
    numpy.random.seed(1)
    dm = numpy.random.random((13, 13))
    dm = dm + dm.T
    vj, vk, _ = oracle.direct_full(h2o_631g, dm[None])
    assert abs(oracle.fp(vk[0]) - pins['pin4_h2o_631g_fp_K_s1il']['value']) < 1e-11
    assert abs(oracle.fp(vj[0]) - pins['pin4_h2o_631g_fp_J_s1kl']['value']) < 1e-11


def test_pin5_complex_dm_veff(oracle, pins):
    mol = gto.M(atom=H2O, basis='631g', cart=True)
    opt = oracle.VHFOpt(mol, 1e-13)
    numpy.random.seed(1)
    dm = numpy.random.random((2, 3, 13, 13)) - .5 + 0j
    vj, vk = oracle.get_jk(mol, dm, hermi=0, vhfopt=opt)
    assert abs(oracle.fp(vj - .5 * vk) - pins['pin5_h2o_631g_cart_fp_vhf4']['value']) < 1e-11


def test_screening_changes_little(oracle, h2o_vdz):
    rng = numpy.random.default_rng(0)
    c = rng.standard_normal((24, 5))
    dm = c @ c.T
    vj0, vk0 = oracle.get_jk(h2o_vdz, dm)
    opt = oracle.VHFOpt(h2o_vdz, 1e-13)
    vj1, vk1 = oracle.get_jk(h2o_vdz, dm, vhfopt=opt)
    assert abs(vj1 - vj0).max() < 1e-11 and abs(vk1 - vk0).max() < 1e-11


def test_pin_energies(oracle, pins, h2o_631g, h2o_vdz):
    assert abs(h2o_631g.energy_nuc() - pins['pin7_h2o_631g_energy_nuc']['value']) < 1e-10
    e, _, _ = oracle.rhf(h2o_631g, conv_tol=1e-11)
    assert abs(e - pins['pin7_h2o_631g_rhf']['value']) < 1e-9
    e, _, _ = oracle.rhf(h2o_vdz, conv_tol=1e-11)
    assert abs(e - pins['pin6_h2o_ccpvdz_rhf_direct']['value']) < 1e-9


def test_pin8_converged_vj(oracle, pins):
    mol = gto.M(atom=H2O, basis='631g', cart=True)
    e, dm, _ = oracle.rhf(mol, conv_tol=1e-10)
    vj, _ = oracle.get_jk(mol, dm)
    assert abs(oracle.fp(vj) - pins['pin8_h2o_631g_cart_fp_vj_converged']['value']) < 1e-6


def test_planted_mistake_is_caught(oracle, pins, h2o_vdz):
    # dropping half of the K permutations (a plausible digestion bug) moves
    # the pinned norm far outside its 1e-9 tolerance
    numpy.random.seed(1)
    dm = numpy.random.random((24, 24))
    vj, vk = oracle.get_jk(h2o_vdz, dm, hermi=0)
    wrong = numpy.tril(vk) + numpy.tril(vk, -1).T
    assert abs(numpy.linalg.norm(wrong) - pins['pin2_h2o_ccpvdz_norm_K']['value']) > 1e-3


@pytest.mark.parametrize('n', [1, 2, 3, 5, 7, 9])
def test_interpolated_roots_match_boys_moments(oracle, n):
    """The oracle's interpolated roots (set_roots(1), used for full-size
    molecules) reproduce the Boys moments F_m(T), m < 2n, like the per-T
    roots; T includes grid nodes, midpoints and the Hermite switch."""
    rng = numpy.random.default_rng(100 + n)
    Ts = numpy.concatenate([[0.0, 1 / 64, 1 / 32, 5.0 + 1 / 64, 36.0 + 6 * n - 1e-9],
                            rng.uniform(0, 36.0 + 6 * n, 30)])
    oracle.set_roots(1)
    try:
        for T in Ts:
            x, w = oracle.rys_roots(n, float(T))   # per-T roots (orc_rys_roots)
            xf = numpy.zeros(n)
            wf = numpy.zeros(n)
            oracle.lib().orc_rys_roots_fast(ctypes.c_int(n), ctypes.c_double(T), xf.ctypes.data_as(ctypes.c_void_p),
                                            wf.ctypes.data_as(ctypes.c_void_p))
            assert abs(xf - x).max() < 1e-13 and (abs(wf - w) / w).max() < 1e-13, (n, T)
            for m in range(2 * n):
                ref = float(oracle.boys_mp(m, T))
                assert abs(numpy.dot(wf, xf**m) - ref) <= 2e-13 * ref, (n, T, m)
    finally:
        oracle.set_roots(0)


def test_interpolated_roots_same_jk(oracle, h2o_vdz):
    """J/K of H2O cc-pVDZ with the interpolated roots equal the per-T ones."""
    numpy.random.seed(1)
    dm = numpy.random.random((24, 24))
    j0, k0 = oracle.get_jk(h2o_vdz, dm, hermi=0)
    oracle.set_roots(1)
    try:
        j1, k1 = oracle.get_jk(h2o_vdz, dm, hermi=0)
    finally:
        oracle.set_roots(0)
    assert abs(j1 - j0).max() < 1e-12 and abs(k1 - k0).max() < 1e-12
