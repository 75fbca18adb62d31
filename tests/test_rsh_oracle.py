"""Range-separated Coulomb operators in the CPU oracle (env[PTR_RANGE_OMEGA]).

omega > 0 is erf(omega r)/r, omega < 0 is erfc(|omega| r)/r
(pyscf/gto/mole.py:2930-2979).  The oracle's Rys evaluation of them is pinned
by the textbook (ss|ss) formula with erf, by a composite Gauss-Legendre
quadrature of the t-integral over the operator's own interval (no Rys roots,
no root scaling, no erfc = 1 - erf split), by erf + erfc = 1 on whole ERI
tensors and by the omega -> infinity limit."""
import math

import numpy
import pytest

from pyscf_amd import gto
from tests.conftest import H2O

OMEGAS = [0.4, -0.4, 3.0, -3.0]


@pytest.fixture
def rys_mode(oracle):
    oracle.set_quadrature(0)
    yield oracle
    oracle.set_quadrature(0)


def _ss_closed(a, b, A, B, omega, s0=0.282094791773878143):
    """(ab|ab)-type primitive (s_a(A) s_a(A) | s_b(B) s_b(B)) from
    F_0(x) = sqrt(pi/x)/2 erf(sqrt(x)): the erf kernel replaces T by
    theta T and multiplies by sqrt(theta), theta = omega^2/(omega^2 + rho)."""
    p, q = 2 * a, 2 * b
    rho = p * q / (p + q)
    T = rho * float(((A - B) ** 2).sum())

    def norm(e):   # gto_norm(0, e) (pyscf/gto/mole.py:125-155) times the s angular factor
        return s0 / math.sqrt(math.gamma(1.5) / (2 * (2 * e) ** 1.5))

    def f0(x):
        return 1.0 if x == 0 else 0.5 * math.sqrt(math.pi / x) * math.erf(math.sqrt(x))

    pref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * norm(a) ** 2 * norm(b) ** 2
    full = pref * f0(T)
    if omega == 0:
        return full
    th = omega ** 2 / (omega ** 2 + rho)
    lr = pref * math.sqrt(th) * f0(th * T)
    return lr if omega > 0 else full - lr


@pytest.mark.parametrize('mode', [0, 1])
@pytest.mark.parametrize('omega', [0.0] + OMEGAS)
def test_ssss_erf_closed_form(oracle, mode, omega):
    a, b = 1.3, 0.45
    A, B = numpy.zeros(3), numpy.array([0.3, -0.2, 1.1])
    mol = gto.M(atom='H 0 0 0; He 0.3 -0.2 1.1', unit='Bohr',
                basis={'H': [[0, [a, 1.0]]], 'He': [[0, [b, 1.0]]]})
    try:
        oracle.set_quadrature(mode, 8)
        with mol.with_range_coulomb(omega):
            v = oracle.int2e_shell(mol, (0, 0, 1, 1)).ravel()[0]
    finally:
        oracle.set_quadrature(0)
    ref = _ss_closed(a, b, A, B, omega)
    assert abs(v - ref) <= 2e-14 * abs(ref) + 1e-16, (v, ref)
    assert mol.omega == 0.0     # the context restored the operator


@pytest.mark.parametrize('omega', [0.35, -0.35, 2.0, -2.0])
def test_rys_roots_vs_t_quadrature_df(rys_mode, omega):
    """Shell quartets with d and f shells: Rys roots (scaled, two sets for
    erfc) against composite Gauss-Legendre over t in [0, sqrt(theta)] (erf)
    or [sqrt(theta), 1] (erfc)."""
    oracle = rys_mode
    mol = gto.M(atom=H2O, basis='def2-tzvp')
    ls = [mol.bas_angular(i) for i in range(mol.nbas)]
    f = ls.index(3)
    d = ls.index(2)
    rng = numpy.random.default_rng(3)
    quartets = [tuple(int(x) for x in rng.integers(0, mol.nbas, 4)) for _ in range(40)]
    quartets += [(f, d, f, 0), (f, f, d, d), (f, f, f, f), (d, 5, f, 1)]
    scale = 0.0
    worst = 0.0
    with mol.with_range_coulomb(omega):
        for qt in quartets:
            oracle.set_quadrature(0)
            a = oracle.int2e_shell(mol, qt)
            oracle.set_quadrature(1, 16)
            b = oracle.int2e_shell(mol, qt)
            scale = max(scale, abs(b).max())
            worst = max(worst, abs(a - b).max())
    assert worst <= 1e-13 * scale, (worst, scale)


def test_erf_plus_erfc_is_coulomb(rys_mode, h2o_631g):
    oracle = rys_mode
    full = oracle.int2e_full(h2o_631g)
    with h2o_631g.with_range_coulomb(0.5):
        lr = oracle.int2e_full(h2o_631g)
    with h2o_631g.with_range_coulomb(-0.5):
        sr = oracle.int2e_full(h2o_631g)
    assert abs(lr + sr - full).max() < 1e-13
    assert abs(lr).max() > 0.1 and abs(sr).max() > 0.1     # both parts are non-trivial
    # long range is bounded by the full operator on the (ij|ij) diagonal
    n = full.shape[0]
    dg = numpy.einsum('ijij->ij', full)
    assert numpy.all(numpy.einsum('ijij->ij', lr) <= dg + 1e-14)
    assert n == 13


def test_large_omega_limit(rys_mode, h2o_631g):
    oracle = rys_mode
    full = oracle.int2e_full(h2o_631g)
    with h2o_631g.with_range_coulomb(1e5):
        lr = oracle.int2e_full(h2o_631g)
    with h2o_631g.with_range_coulomb(-1e5):
        sr = oracle.int2e_full(h2o_631g)
    assert abs(lr - full).max() < 1e-7 * abs(full).max()
    assert abs(sr).max() < 1e-7 * abs(full).max()


def test_jk_linear_in_operator(rys_mode, h2o_vdz):
    """J/K of erf and erfc add up to the Coulomb J/K (oracle get_jk)."""
    oracle = rys_mode
    rng = numpy.random.default_rng(7)
    dm = rng.standard_normal((24, 24))
    j0, k0 = oracle.get_jk(h2o_vdz, dm, hermi=0)
    with h2o_vdz.with_range_coulomb(0.3):
        j1, k1 = oracle.get_jk(h2o_vdz, dm, hermi=0)
    with h2o_vdz.with_range_coulomb(-0.3):
        j2, k2 = oracle.get_jk(h2o_vdz, dm, hermi=0)
    assert abs(j1 + j2 - j0).max() < 1e-12 and abs(k1 + k2 - k0).max() < 1e-12
    assert abs(j1 - j0).max() > 1e-2
