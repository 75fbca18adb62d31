"""Closed-shell RHF driver that calls the GPU J/K engine unchanged.

Restates the reference SCF loop (pyscf/scf/hf.py:47-242):
  * fock = h + vhf, CDIIS extrapolation on the commutator S D F - F D S
    (pyscf/scf/diis.py:40-66, pyscf/lib/diis.py:202-275), from cycle 1;
  * generalized eigenproblem F C = S C e (pyscf/scf/hf.py:1340-1347);
  * aufbau occupation, D = 2 C_occ C_occ^T (pyscf/scf/hf.py:840);
  * incremental Fock build vhf = vhf_last + J(dD) - K(dD)/2 with screening
    at direct_scf_tol (pyscf/scf/hf.py:2416-2427);
  * optional level shift of the virtual space, F + shift (S - S D/2 S),
    applied after DIIS (pyscf/scf/hf.py:766-786, 1123-1124);
  * converged when |dE| < conv_tol and |g| < sqrt(conv_tol)
    (pyscf/scf/hf.py:194-197).
The initial guess is the core Hamiltonian ('1e'); the reference defaults to
'minao', which changes the iteration path but not the converged energy.
One-electron integrals come from the library's host code (mjk_int1e).
"""

import ctypes
import time

import numpy
import scipy.linalg

from pyscf_amd import lib as _lib
from pyscf_amd.scf import _vhf, hf


def int1e(mol, which):
    """'ovlp', 'kin' or 'nuc' matrix (spherical AOs)."""
    l = _lib.load_library()
    code = {'ovlp': 0, 'kin': 1, 'nuc': 2}[which]
    atm = numpy.ascontiguousarray(mol._atm, dtype=numpy.int32)
    bas = numpy.ascontiguousarray(mol._bas, dtype=numpy.int32)
    env = numpy.ascontiguousarray(mol._env, dtype=numpy.double)
    nao = int(mol.nao_nr())
    out = numpy.zeros((nao, nao))
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = l.mjk_int1e(code, p(atm), len(atm), p(bas), len(bas), p(env), p(out))
    if rc != 0:
        raise RuntimeError('mjk_int1e failed (%d)' % rc)
    return out


def int1e_device(mol, engine=None, which=('ovlp', 'kin', 'nuc')):
    """The same matrices computed on the GPU (mjk_int1e_device), as a dict."""
    eng = engine or _vhf.get_engine()
    atm = numpy.ascontiguousarray(mol._atm, dtype=numpy.int32)
    bas = numpy.ascontiguousarray(mol._bas, dtype=numpy.int32)
    env = numpy.ascontiguousarray(mol._env, dtype=numpy.double)
    nao = int(mol.nao_nr())
    out = {w: numpy.empty((nao, nao)) for w in which}
    p = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
    eng.check(eng.lib.mjk_int1e_device(eng.handle, p(atm), len(atm), p(bas), len(bas), p(env), p(out.get('ovlp')),
                                       p(out.get('kin')), p(out.get('nuc')), 0), 'mjk_int1e_device')
    return out


class CDIIS:
    """Commutator DIIS on the Fock matrix."""

    def __init__(self, space=8):
        self.space = space
        self.fs = []
        self.es = []

    def update(self, s, d, f):
        sdf = s @ d @ f
        self.fs.append(f)
        self.es.append(sdf.T - sdf)
        self.fs = self.fs[-self.space:]
        self.es = self.es[-self.space:]
        n = len(self.fs)
        if n < 2:
            return f
        B = numpy.empty((n + 1, n + 1))
        B[0, 0] = 0
        B[0, 1:] = B[1:, 0] = 1
        for i in range(n):
            for j in range(i + 1):
                B[i + 1, j + 1] = B[j + 1, i + 1] = numpy.vdot(self.es[i], self.es[j])
        rhs = numpy.zeros(n + 1)
        rhs[0] = 1
        try:
            c = scipy.linalg.solve(B, rhs, assume_a='sym')[1:]
        except (numpy.linalg.LinAlgError, ValueError):
            c = numpy.linalg.lstsq(B, rhs, rcond=None)[0][1:]
        return sum(ci * fi for ci, fi in zip(c, self.fs))


class RHF:
    def __init__(self, mol, direct_scf_tol=1e-13, fp32_tol=0.0):
        self.mol = mol
        self.conv_tol = 1e-9
        self.conv_tol_grad = None
        self.max_cycle = 100
        self.direct_scf_tol = direct_scf_tol
        self.fp32_tol = fp32_tol
        self.diis_space = 8
        self.level_shift = 0.0
        self.verbose = 0
        self._opt = None
        self.e_tot = 0.0
        self.converged = False
        self.cycles = 0
        self.jk_times = []
        self.mo_energy = self.mo_coeff = self.mo_occ = None

    # -- integrals -------------------------------------------------------
    def get_hcore(self, mol=None):
        mol = mol or self.mol
        return int1e(mol, 'kin') + int1e(mol, 'nuc')

    def get_ovlp(self, mol=None):
        return int1e(mol or self.mol, 'ovlp')

    def init_direct_scf(self, mol=None):
        return _vhf.VHFOpt(mol or self.mol, direct_scf_tol=self.direct_scf_tol)

    def get_jk(self, mol=None, dm=None, hermi=1, with_j=True, with_k=True, omega=None):
        """One screening context per omega, built under that operator
        (pyscf/scf/hf.py:2057-2080)."""
        mol = mol or self.mol
        if self._opt is None:
            self._opt = {}
        if self._opt.get(omega) is None:
            with mol.with_range_coulomb(omega):
                self._opt[omega] = self.init_direct_scf(mol)
        t0 = time.perf_counter()
        vj, vk = hf.get_jk(mol, dm, hermi, self._opt[omega], with_j, with_k, omega, fp32_tol=self.fp32_tol)
        self.jk_times.append(time.perf_counter() - t0)
        return vj, vk

    def get_veff(self, mol=None, dm=None, dm_last=None, vhf_last=None, hermi=1):
        if dm_last is None or vhf_last is None:
            vj, vk = self.get_jk(mol, dm, hermi)
            return vj - vk * .5
        vj, vk = self.get_jk(mol, numpy.asarray(dm) - numpy.asarray(dm_last), hermi)
        return vj - vk * .5 + vhf_last

    # -- SCF pieces --------------------------------------------------------
    def eig(self, f, s):
        """F C = S C e (pyscf/scf/hf.py:1340-1347), through the symmetric
        orthogonaliser X = S^-1/2 of the overlap, computed once per S: a
        standard eigenproblem of X^T F X per cycle (3x faster than the
        generalized solver at 400 AOs).  Nearly singular overlaps (smallest
        eigenvalue below 1e-10) keep the generalized solver."""
        cache = getattr(self, '_orth', None)
        if cache is None or cache[0] is not s:
            w, v = numpy.linalg.eigh(s)
            cache = (s, v / numpy.sqrt(w) if w[0] > 1e-10 else None)
            self._orth = cache
        x = cache[1]
        if x is None:
            return scipy.linalg.eigh(f, s)
        e, c = numpy.linalg.eigh(x.T @ f @ x)
        return e, x @ c

    def density_fit(self, auxbasis=None, with_df=None):
        """This SCF with density-fitted J/K (pyscf/df/df_jk.py:32-125)."""
        from pyscf_amd.df import df_jk
        return df_jk.density_fit(self, auxbasis, with_df)

    def get_occ(self, mo_energy):
        nocc = self.mol.nelectron // 2
        occ = numpy.zeros_like(mo_energy)
        occ[numpy.argsort(mo_energy)[:nocc]] = 2
        return occ

    @staticmethod
    def make_rdm1(mo_coeff, mo_occ):
        c = mo_coeff[:, mo_occ > 0]
        return (c * mo_occ[mo_occ > 0]) @ c.T

    @staticmethod
    def energy_elec(dm, h1e, vhf):
        e1 = numpy.einsum('ij,ji->', h1e, dm)
        e2 = .5 * numpy.einsum('ij,ji->', vhf, dm)
        return e1 + e2

    def get_grad(self, mo_coeff, mo_occ, fock):
        occ = mo_occ > 0
        return (mo_coeff[:, ~occ].T @ fock @ mo_coeff[:, occ]) * 2

    def kernel(self, dm0=None):
        mol = self.mol
        if mol.nelectron % 2:
            raise ValueError('RHF needs an even number of electrons')
        h1e = self.get_hcore()
        s1e = self.get_ovlp()
        enuc = mol.energy_nuc()
        conv_tol_grad = self.conv_tol_grad or numpy.sqrt(self.conv_tol)
        if dm0 is None:
            e, c = self.eig(h1e, s1e)
            dm = self.make_rdm1(c, self.get_occ(e))
        else:
            dm = dm0
        vhf = self.get_veff(mol, dm)
        e_tot = self.energy_elec(dm, h1e, vhf) + enuc
        diis = CDIIS(self.diis_space)
        self.converged = False
        for cycle in range(self.max_cycle):
            dm_last, last_e = dm, e_tot
            fock = h1e + vhf
            fock = diis.update(s1e, dm, fock)
            if abs(self.level_shift) > 1e-4:
                fock = fock + self.level_shift * (s1e - s1e @ (dm * .5) @ s1e)
            mo_energy, mo_coeff = self.eig(fock, s1e)
            mo_occ = self.get_occ(mo_energy)
            dm = self.make_rdm1(mo_coeff, mo_occ)
            vhf = self.get_veff(mol, dm, dm_last, vhf)
            e_tot = self.energy_elec(dm, h1e, vhf) + enuc
            fock = h1e + vhf
            norm_g = numpy.linalg.norm(self.get_grad(mo_coeff, mo_occ, fock))
            if self.verbose:
                print('cycle %d E=%.15g dE=%.3g |g|=%.3g' % (cycle + 1, e_tot, e_tot - last_e, norm_g))
            self.cycles = cycle + 1
            if abs(e_tot - last_e) < self.conv_tol and norm_g < conv_tol_grad:
                self.converged = True
                break
        self.e_tot = e_tot
        self.mo_energy, self.mo_coeff, self.mo_occ = mo_energy, mo_coeff, mo_occ
        self.dm = dm
        return e_tot

    scf = kernel

    def make_rdm1_final(self):
        return self.make_rdm1(self.mo_coeff, self.mo_occ)


class DeviceRHF(RHF):
    """RHF whose whole loop runs on the device (mjk_rhf_create/run/fetch):
    the density, Fock matrix, DIIS history and MO coefficients stay in HBM,
    the dense algebra runs on rocBLAS/rocSOLVER, J/K on the engine.  Same
    inputs, attributes and convergence test as RHF; the host sees the DIIS
    overlaps and three scalars per cycle.  The initial guess is the core
    Hamiltonian, or dm0.  With ``with_df`` (``density_fit()``) every cycle
    builds density-fitted J/K of the full density, K from the occupied
    orbitals, all on the device."""

    def kernel(self, dm0=None):
        mol = self.mol
        if mol.nelectron % 2:
            raise ValueError('RHF needs an even number of electrons')
        enuc = float(mol.energy_nuc())
        conv_tol_grad = self.conv_tol_grad or numpy.sqrt(self.conv_tol)
        if self._opt is None:
            self._opt = {}
        if self._opt.get(None) is None:
            self._opt[None] = self.init_direct_scf(mol)
        opt = self._opt[None]
        ints = int1e_device(mol, opt.engine)      # one-electron integrals on the device too
        h1e = numpy.ascontiguousarray(ints['kin'] + ints['nuc'])
        s1e = ints['ovlp']
        eng, L = opt.engine, opt.engine.lib
        nao = s1e.shape[0]
        p = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
        h = ctypes.c_void_p()
        eng.check(L.mjk_rhf_create(eng.handle, opt.basis.handle, p(s1e), p(h1e), mol.nelectron // 2, enuc,
                                   self.diis_space, ctypes.byref(h)), 'mjk_rhf_create')
        try:
            with_df = getattr(self, 'with_df', None)
            if with_df is not None:
                if with_df.engine is None:
                    with_df.engine = eng
                if with_df.engine is not eng:
                    raise ValueError('the density fit belongs to another engine')
                with_df.build()
                eng.check(L.mjk_rhf_set_df(eng.handle, h, with_df.handle), 'mjk_rhf_set_df')
            if dm0 is not None:
                dm0 = numpy.ascontiguousarray(dm0, dtype=numpy.double)
            res = _lib.mjk_rhf_result()
            t0 = time.perf_counter()
            eng.check(L.mjk_rhf_run(eng.handle, h, p(dm0), self.max_cycle, self.conv_tol, conv_tol_grad,
                                    self.direct_scf_tol, self.fp32_tol, self.level_shift, ctypes.byref(res)),
                      'mjk_rhf_run')
            self.seconds = time.perf_counter() - t0
            dm = numpy.empty((nao, nao))
            mo_energy = numpy.empty(nao)
            mo_coeff = numpy.empty((nao, nao))
            eng.check(L.mjk_rhf_fetch(eng.handle, h, p(dm), p(mo_energy), p(mo_coeff), None), 'mjk_rhf_fetch')
        finally:
            L.mjk_rhf_destroy(h)
        self.e_tot = res.e_tot
        self.converged = bool(res.converged)
        self.cycles = res.cycles
        self.gnorm = res.gnorm
        self.mo_energy, self.mo_coeff = mo_energy, mo_coeff
        self.mo_occ = self.get_occ(mo_energy)
        self.dm = dm
        return self.e_tot

    scf = kernel
