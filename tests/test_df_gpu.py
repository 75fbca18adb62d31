"""Density-fitted J/K on the GPU (mjk_df_*) against the oracle and the
reference's golden values.

* The raw integrals (a build with lindep < 0 keeps (ij|P)) and (P|Q) equal
  the oracle's to 1e-12 relative (same Rys formulation, different code).
* The fitted tensor and J/K of random densities equal the oracle's fit to
  the FP64 contract (1e-10), for H2O cc-pVDZ / weigend (g auxiliary
  functions), a cc-pVDZ chain with generally contracted s shells /
  cc-pVDZ-JKFIT, and a def2-TZVP chain (f orbital shells) /
  def2-universal-JKFIT (g).
* Golden values of pyscf/df/test/test_df_jk.py: lib.fp(vj), lib.fp(vk) of
  the seed-1 densities (9 places) and the DF-RHF energy (8 places).
* K from occupied orbitals (the SCF path) equals K of their density.
* The exchange GEMMs on even dimensions and in lower blocks (odd AO count,
  forced blocks) against the oracle.
"""
import numpy
import pytest

from pyscf_amd import df, gto, lib, workloads
from pyscf_amd.scf import rhf

pytestmark = pytest.mark.gpu
H2O = 'O 0 0 0; H 0 -0.757 0.587; H 0 0.757 0.587'
TOL = 1e-10

CASES = {
    'h2o_weigend': (lambda: gto.M(atom=H2O, basis='cc-pvdz'), 'weigend'),
    'chain_vdz_jkfit': (lambda: gto.M(atom=workloads.random_chain([8, 1], nmin=3, nmax=3), basis='cc-pvdz'),
                        None),
    'chain_tzvp_jkfit': (lambda: gto.M(atom=workloads.random_chain([8, 2], nmin=2, nmax=2), basis='def2-tzvp'),
                         None),
}


def dens(nao, seed, n=2, sym=False):
    d = numpy.random.default_rng(seed).standard_normal((n, nao, nao)) / nao
    return d + d.transpose(0, 2, 1) if sym else d


@pytest.mark.parametrize('name', sorted(CASES))
def test_raw_integrals(oracle, name):
    mk, auxbasis = CASES[name]
    mol = mk()
    d = df.DF(mol, auxbasis)
    d.lindep = -1.0      # keep the raw (ij|P)
    d.build()
    aux = d.auxmol
    raw = d.cderi()      # (naux, nao, nao) = (ij|P) transposed
    j3 = oracle.int3c2e(mol, aux)
    ref = j3.transpose(2, 0, 1)
    scale = abs(ref).max()
    print(name, mol.nao_nr(), aux.nao_nr(), 'max|d3c|=%.2e (max %.2e)' % (abs(raw - ref).max(), scale), flush=True)
    assert abs(raw - ref).max() <= 1e-12 * scale
    j2 = numpy.empty((aux.nao_nr(), aux.nao_nr()))
    d.engine.check(d.engine.lib.mjk_df_fetch(d.engine.handle, d.handle, 0, j2.ctypes.data_as(lib.ctypes.c_void_p)),
                   'mjk_df_fetch')
    j2o = oracle.int2c2e(mol, aux)
    assert abs(j2 - j2o).max() <= 1e-12 * abs(j2o).max()


@pytest.mark.parametrize('name', sorted(CASES))
def test_fit_and_jk_vs_oracle(oracle, name):
    mk, auxbasis = CASES[name]
    mol = mk()
    nao = mol.nao_nr()
    d = df.DF(mol, auxbasis).build()
    cd = oracle.cholesky_eri(mol, d.auxmol)
    b = d.cderi()
    assert b.shape == cd.shape
    print(name, 'max|dB|=%.2e' % abs(b - cd).max(), flush=True)
    assert abs(b - cd).max() <= TOL
    for hermi, sym in ((0, False), (1, True)):
        dms = dens(nao, 7 + hermi, sym=sym)
        vj, vk = df.get_jk(d, dms, hermi)
        j0, k0 = oracle.df_get_jk(cd, dms, hermi)
        print('  hermi=%d max|dJ|=%.2e max|dK|=%.2e' % (hermi, abs(vj - j0).max(), abs(vk - k0).max()), flush=True)
        assert abs(vj - j0).max() <= TOL and abs(vk - k0).max() <= TOL


@pytest.mark.parametrize('kblocks', ['0', '3'])
def test_exchange_even_gemms_and_blocks(oracle, monkeypatch, kblocks):
    """The exchange products run their GEMMs on an even number of AOs (an odd
    last AO by matrix-vector products) and, with more than one block, as the
    lower blocks of the symmetric product, mirrored: NH3 cc-pVDZ (29 AOs) and
    H2O cc-pVDZ (24 AOs), general and occupied-orbital paths, against the
    oracle's fit (kblocks 0 keeps one block at these sizes, 3 forces three)."""
    monkeypatch.setenv('MJK_DF_KBLOCKS', kblocks)   # read when a DF object is built
    nh3 = 'N 0 0 0.11; H 0 0.94 -0.26; H 0.81 -0.47 -0.26; H -0.81 -0.47 -0.26'
    for atom in (nh3, H2O):
        mol = gto.M(atom=atom, basis='cc-pvdz')
        nao = mol.nao_nr()
        d = df.DF(mol, 'weigend').build()
        cd = oracle.cholesky_eri(mol, d.auxmol)
        for hermi, sym in ((0, False), (1, True)):
            dms = dens(nao, 11 + hermi, sym=sym)
            vj, vk = df.get_jk(d, dms, hermi)
            j0, k0 = oracle.df_get_jk(cd, dms, hermi)
            print(nao, 'kblocks', kblocks, 'hermi=%d max|dJ|=%.2e max|dK|=%.2e'
                  % (hermi, abs(vj - j0).max(), abs(vk - k0).max()), flush=True)
            assert abs(vj - j0).max() <= TOL and abs(vk - k0).max() <= TOL
        rng = numpy.random.default_rng(3)
        c = rng.standard_normal((nao, nao)) / nao
        occ = numpy.zeros(nao)
        occ[:5] = 2.0
        dm = 2.0 * c[:, :5] @ c[:, :5].T
        vj, vk = df.get_jk(d, lib.tag_array(dm, mo_coeff=c, mo_occ=occ), 1)
        j0, k0 = oracle.df_get_jk(cd, dm[None], 1)
        print(nao, 'kblocks', kblocks, 'orbitals max|dK|=%.2e' % abs(vk - k0[0]).max(), flush=True)
        assert abs(vj - j0[0]).max() <= TOL and abs(vk - k0[0]).max() <= TOL


def test_golden_fingerprints():
    mol = gto.M(atom=H2O, basis='cc-pvdz')
    d = df.DF(mol, 'weigend')
    dms = numpy.random.RandomState(1).random_sample((2, 24, 24))
    vj, vk = df.get_jk(d, dms, hermi=0)
    print('fp(vj)=%.12f fp(vk)=%.12f' % (lib.fp(vj), lib.fp(vk)), flush=True)
    assert abs(lib.fp(vj) - -194.15910890730066) < 1e-9
    assert abs(lib.fp(vk) - -46.365071587653517) < 1e-9


def test_df_rhf_golden_and_orbital_path():
    mol = gto.M(atom=H2O, basis='cc-pvdz')
    mf = rhf.RHF(mol).density_fit(auxbasis='weigend')
    mf.conv_tol = 1e-11
    e = mf.kernel()
    print('DF-RHF E=%.12f (%d cycles)' % (e, mf.cycles), flush=True)
    assert mf.converged and abs(e - -76.025936299702536) < 1e-8
    # K from the occupied orbitals equals K of their density
    dm = mf.make_rdm1(mf.mo_coeff, mf.mo_occ)
    vj1, vk1 = df.get_jk(mf.with_df, dm, 1)
    vj2, vk2 = df.get_jk(mf.with_df, numpy.asarray(dm).copy(), 1)
    assert abs(vj1 - vj2).max() <= 1e-12 and abs(vk1 - vk2).max() <= 1e-11


def test_cderi_packed_layout():
    mol = gto.M(atom=H2O, basis='cc-pvdz')
    d = df.DF(mol, 'weigend').build()
    b = d.cderi()
    packed = d._cderi
    i, j = numpy.tril_indices(24)
    assert packed.shape == (d.get_naoaux(), 24 * 25 // 2)
    assert abs(packed - b[:, i, j]).max() == 0 and abs(b - b.transpose(0, 2, 1)).max() <= 1e-13 * abs(b).max()


def test_device_resident_df_rhf():
    """The density-fitted RHF with the whole loop on the device (mjk_rhf_set_df)
    against the host-driven density-fitted RHF and the reference's golden
    energy."""
    mol = gto.M(atom=H2O, basis='cc-pvdz')
    dev = rhf.DeviceRHF(mol).density_fit(auxbasis='weigend')
    dev.conv_tol = 1e-11
    e = dev.kernel()
    print('device DF-RHF E=%.12f (%d cycles)' % (e, dev.cycles), flush=True)
    assert dev.converged and abs(e - -76.025936299702536) < 1e-8
    mol = CASES['chain_vdz_jkfit'][0]()
    host = rhf.RHF(mol).density_fit()
    host.conv_tol = 1e-11
    eh = host.kernel()
    dev = rhf.DeviceRHF(mol).density_fit()
    dev.conv_tol = 1e-11
    ed = dev.kernel()
    print('chain DF-RHF host %.12f device %.12f' % (eh, ed), flush=True)
    assert host.converged and dev.converged and abs(eh - ed) < 1e-9
