"""ORACLE — TEST INFRASTRUCTURE ONLY.

CPU restatement of PySCF's direct J/K path, used by tests/, by
``__graft_entry__.smoke()`` and by bench.py's ``cpu_baseline`` leg as the
checker / CPU baseline.  The product package ``pyscf_amd`` never imports it.

Python side of the oracle:
  * ``get_jk``       mirrors pyscf/scf/hf.py:948-1022 + pyscf/scf/_vhf.py:405-464
                     (complex split, J/K lower-triangle + mirror per hermi)
  * ``VHFOpt``       q_cond / dm_cond / predicate of pyscf/scf/_vhf.py:151-264
                     and lib/vhf/optimizer.c:90-117, 408-454, 494-518
  * ``hermi_triu``   pyscf/lib/numpy_helper.py:480-525
  * ``fp``           pyscf/lib/misc.py:1259-1263
  * ``rhf``          the RHF SCF loop of pyscf/scf/hf.py:47-242 with CDIIS
                     (pyscf/scf/diis.py:40-66) and the incremental Fock build
                     (pyscf/scf/hf.py:2416-2427); core-Hamiltonian initial guess
  * ``set_roots``    per-T roots (default) or interpolated roots (full-size
                     molecules)
  * ``boys_mp``      Boys function F_m(T) to 30 digits with mpmath (pins the
                     Rys roots: sum_i w_i x_i^m == F_m(T))
  * ``set_quadrature`` the t-integral of every primitive quartet by Rys
                     roots (mode 0) or by composite Gauss-Legendre over the
                     operator's interval (mode 1, a check of mode 0 that
                     shares none of its root machinery).  The operator is
                     env[PTR_RANGE_OMEGA] (8), as in libcint: 0 Coulomb,
                     > 0 erf(omega r)/r, < 0 erfc(|omega| r)/r.
The ERI engine (Rys quadrature, Stieltjes/Golub-Welsch roots) and the s8
shell-quartet loop are in ``jk_oracle.c`` beside this file.
"""

import ctypes
import os
import subprocess

import numpy

_DIR = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_DIR, 'liborc_jk.so')
_lib = None


def build():
    subprocess.check_call(['make', '-s', '-C', _DIR])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            build()
        _lib = ctypes.CDLL(_LIBPATH)
        _lib.orc_init()
        _lib.orc_direct_jk.restype = ctypes.c_long
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _mol_arrays(mol):
    atm = numpy.asarray(mol._atm, dtype=numpy.int32, order='C')
    bas = numpy.asarray(mol._bas, dtype=numpy.int32, order='C')
    env = numpy.asarray(mol._env, dtype=numpy.double, order='C')
    l = bas[:, 1]
    ao_loc = numpy.zeros(len(bas) + 1, dtype=numpy.int32)
    ao_loc[1:] = numpy.cumsum((2 * l + 1) * bas[:, 3])
    return atm, bas, env, ao_loc


def set_threads(n):
    lib().orc_set_threads(ctypes.c_int(n))


def fp(a):
    """lib.fp: cos(arange(n)) . a.ravel() (pyscf/lib/misc.py:1259-1263)."""
    a = numpy.asarray(a)
    return numpy.dot(numpy.cos(numpy.arange(a.size)), a.ravel())


def hermi_triu(mat, hermi=1):
    """Fill the upper triangle from the lower one (hermi=1: mirror,
    hermi=2: anti-mirror), pyscf/lib/numpy_helper.py:480-525."""
    mat = mat.copy()
    n = mat.shape[-1]
    iu = numpy.triu_indices(n, 1)
    if hermi == 1:
        mat[..., iu[0], iu[1]] = mat[..., iu[1], iu[0]]
    elif hermi == 2:
        mat[..., iu[0], iu[1]] = -mat[..., iu[1], iu[0]]
    return mat


def set_roots(mode):
    """0: roots by the per-T Stieltjes procedure (default); 1: interpolated
    from that procedure's values on a 1/32 grid in T (8-point Lagrange,
    ~1e-14 relative; for full-size molecules)."""
    lib().orc_set_roots(ctypes.c_int(mode))


def set_quadrature(mode, panels=8):
    lib().orc_set_quadrature(ctypes.c_int(mode), ctypes.c_int(panels))


def rys_roots(n, T):
    x = numpy.zeros(n)
    w = numpy.zeros(n)
    lib().orc_rys_roots(ctypes.c_int(n), ctypes.c_double(T), _p(x), _p(w))
    return x, w


def boys_mp(m, T, dps=30):
    import mpmath as mp
    with mp.workdps(dps):
        T = mp.mpf(T)
        if T == 0:
            return mp.mpf(1) / (2 * m + 1)
        return mp.gammainc(m + mp.mpf(1) / 2, 0, T) / (2 * T ** (m + mp.mpf(1) / 2))


def int2e_full(mol):
    atm, bas, env, ao_loc = _mol_arrays(mol)
    nao = int(ao_loc[-1])
    eri = numpy.zeros((nao,) * 4)
    lib().orc_int2e_full(_p(eri), _p(ao_loc), _p(atm), ctypes.c_int(len(atm)),
                         _p(bas), ctypes.c_int(len(bas)), _p(env))
    return eri


def int2e_shell(mol, shls):
    atm, bas, env, ao_loc = _mol_arrays(mol)
    d = [int(ao_loc[s + 1] - ao_loc[s]) for s in shls]
    out = numpy.zeros(d[0] * d[1] * d[2] * d[3])
    sh = numpy.asarray(shls, dtype=numpy.int32)
    lib().orc_int2e_sph(_p(out), _p(sh), _p(atm), ctypes.c_int(len(atm)),
                        _p(bas), ctypes.c_int(len(bas)), _p(env))
    return out.reshape(d[::-1]).transpose(3, 2, 1, 0)


def pack_s8(eri):
    """s8-packed ERI as mol.intor('int2e', aosym='s8')."""
    nao = eri.shape[0]
    tril = numpy.tril_indices(nao)
    eri2 = eri[tril[0], tril[1]][:, tril[0], tril[1]]
    npair = eri2.shape[0]
    return eri2[numpy.tril_indices(npair)]


def int1e(mol, which):
    atm, bas, env, ao_loc = _mol_arrays(mol)
    nao = int(ao_loc[-1])
    out = numpy.zeros((nao, nao))
    code = {'ovlp': 0, 'kin': 1, 'nuc': 2}[which]
    lib().orc_int1e(_p(out), ctypes.c_int(code), _p(ao_loc), _p(atm), ctypes.c_int(len(atm)),
                    _p(bas), ctypes.c_int(len(bas)), _p(env))
    return out


_PRESCREEN = {'CVHFnrs8_prescreen': 0, 'CVHFnrs8_vj_prescreen': 1, 'CVHFnrs8_vk_prescreen': 2}


class VHFOpt:
    """Screening context: Schwarz q_cond once per molecule, dm_cond per call;
    ``prescreen`` names the quartet predicate (lib/vhf/optimizer.c:90-167)."""

    def __init__(self, mol, direct_scf_tol=1e-13, prescreen='CVHFnrs8_prescreen'):
        atm, bas, env, ao_loc = _mol_arrays(mol)
        self.mol = mol
        self.direct_scf_tol = direct_scf_tol
        self.prescreen = prescreen
        nbas = len(bas)
        self.q_cond = numpy.zeros((nbas, nbas))
        lib().orc_q_cond(_p(self.q_cond), _p(ao_loc), _p(atm), ctypes.c_int(len(atm)),
                         _p(bas), ctypes.c_int(nbas), _p(env))

    def dm_cond(self, dms):
        atm, bas, env, ao_loc = _mol_arrays(self.mol)
        nbas = len(bas)
        dms = numpy.ascontiguousarray(dms, dtype=numpy.double)
        out = numpy.zeros((nbas, nbas))
        lib().orc_dm_cond(_p(out), _p(dms), ctypes.c_int(dms.shape[0]), _p(ao_loc), ctypes.c_int(nbas))
        return out


def direct_full(mol, dms, vhfopt=None, with_j=True, with_k=True, pair_stride=1, pair_offset=0):
    """Full J and full K (all 8 permutations) for a stack of real dms.
    Returns (vj, vk, n_quartets_evaluated)."""
    atm, bas, env, ao_loc = _mol_arrays(mol)
    dms = numpy.ascontiguousarray(dms, dtype=numpy.double)
    n_dm, nao = dms.shape[0], dms.shape[-1]
    vj = numpy.zeros((n_dm, nao, nao))
    vk = numpy.zeros((n_dm, nao, nao))
    if vhfopt is not None:
        q = numpy.ascontiguousarray(vhfopt.q_cond)
        dc = vhfopt.dm_cond(dms)
        qp, dcp, cut = _p(q), _p(dc), vhfopt.direct_scf_tol
        lib().orc_set_prescreen(ctypes.c_int(_PRESCREEN[vhfopt.prescreen]))
    else:
        qp = dcp = None
        cut = 0.
    nq = lib().orc_direct_jk(_p(dms), ctypes.c_int(n_dm), _p(vj), _p(vk),
                             ctypes.c_int(int(with_j)), ctypes.c_int(int(with_k)),
                             qp, dcp, ctypes.c_double(cut),
                             _p(ao_loc), _p(atm), ctypes.c_int(len(atm)),
                             _p(bas), ctypes.c_int(len(bas)), _p(env),
                             ctypes.c_int(pair_stride), ctypes.c_int(pair_offset))
    lib().orc_set_prescreen(ctypes.c_int(0))
    return vj, vk, nq


def get_jk(mol, dm, hermi=1, vhfopt=None, with_j=True, with_k=True):
    """Oracle of scf.hf.get_jk (pyscf/scf/hf.py:948-1022)."""
    dm = numpy.asarray(dm, order='C')
    shape = dm.shape
    nao = shape[-1]
    cplx = dm.dtype == numpy.complex128
    if cplx:
        dms = numpy.vstack((dm.real.reshape(-1, nao, nao), dm.imag.reshape(-1, nao, nao)))
        hermi = 0
    else:
        dms = dm.reshape(-1, nao, nao).astype(numpy.double)
    vj, vk, _ = direct_full(mol, dms, vhfopt, with_j, with_k)
    # J is computed on the lower triangle and mirrored in the reference; the
    # full J here is already symmetric.  K follows hermi_triu(vk, hermi).
    if hermi != 0:
        vk = hermi_triu(vk, hermi)
    if cplx:
        vj = (vj[:len(vj) // 2] + 1j * vj[len(vj) // 2:]).reshape(shape)
        vk = (vk[:len(vk) // 2] + 1j * vk[len(vk) // 2:]).reshape(shape)
    else:
        vj = vj.reshape(shape)
        vk = vk.reshape(shape)
    return (vj if with_j else None), (vk if with_k else None)


def jk_einsum(eri, dm):
    """dot_eri_dm definition (pyscf/scf/hf.py:891-936)."""
    vj = numpy.einsum('ijkl,...ji->...kl', eri, dm)
    vk = numpy.einsum('ijkl,...jk->...il', eri, dm)
    return vj, vk


def rhf(mol, conv_tol=1e-10, max_cycle=100, direct_scf_tol=1e-13, get_jk_fn=None, diis_space=8):
    """Closed-shell RHF with CDIIS; returns (e_tot, dm, mo_energy)."""
    S = int1e(mol, 'ovlp')
    H = int1e(mol, 'kin') + int1e(mol, 'nuc')
    enuc = mol.energy_nuc()
    nocc = mol.nelectron // 2
    opt = VHFOpt(mol, direct_scf_tol)
    if get_jk_fn is None:
        def get_jk_fn(d):
            return get_jk(mol, d, hermi=1, vhfopt=opt)
    import scipy.linalg

    def eig(F):
        e, c = scipy.linalg.eigh(F, S)
        return e, c

    e, c = eig(H)
    dm = 2 * c[:, :nocc] @ c[:, :nocc].T
    vj, vk = get_jk_fn(dm)
    vhf = vj - .5 * vk
    e_tot = numpy.einsum('ij,ji->', dm, H) + .5 * numpy.einsum('ij,ji->', dm, vhf) + enuc
    fs, es = [], []
    for cycle in range(max_cycle):
        dm_last, e_last = dm, e_tot
        F = H + vhf
        err = F @ dm @ S - S @ dm @ F
        fs.append(F)
        es.append(err)
        fs, es = fs[-diis_space:], es[-diis_space:]
        if len(fs) > 1:
            n = len(fs)
            B = numpy.zeros((n + 1, n + 1))
            B[0, 1:] = B[1:, 0] = -1
            for i in range(n):
                for j in range(n):
                    B[i + 1, j + 1] = numpy.vdot(es[i], es[j])
            rhs = numpy.zeros(n + 1)
            rhs[0] = -1
            coef = numpy.linalg.solve(B, rhs)[1:]
            F = sum(ci * fi for ci, fi in zip(coef, fs))
        e, c = eig(F)
        dm = 2 * c[:, :nocc] @ c[:, :nocc].T
        vj, vk = get_jk_fn(dm - dm_last)
        vhf = vhf + vj - .5 * vk
        e_tot = numpy.einsum('ij,ji->', dm, H) + .5 * numpy.einsum('ij,ji->', dm, vhf) + enuc
        Fn = H + vhf
        g = Fn @ dm @ S - S @ dm @ Fn
        if abs(e_tot - e_last) < conv_tol and numpy.linalg.norm(g) < numpy.sqrt(conv_tol):
            break
    # final full rebuild to remove incremental drift
    vj, vk = get_jk_fn(dm)
    vhf = vj - .5 * vk
    e_tot = numpy.einsum('ij,ji->', dm, H) + .5 * numpy.einsum('ij,ji->', dm, vhf) + enuc
    return e_tot, dm, e


# ---------------- density fitting (pyscf/df/incore.py, pyscf/df/df_jk.py) ----------------

def _df_arrays(mol, auxmol):
    """atm/bas/env with the orbital shells, the auxiliary shells and a unit s
    shell (exponent 0, coefficient 1) appended, as gto.conc_env joins them
    (pyscf/df/incore.py:80-84)."""
    atm, bas, env, ao_loc = _mol_arrays(mol)
    aatm, abas, aenv, aux_loc = _mol_arrays(auxmol)
    off = len(env)
    abas = abas.copy()
    abas[:, 5] += off
    abas[:, 6] += off
    aatm = aatm.copy()
    aatm[:, 1] += off
    aatm[:, 3] += off
    unit = numpy.array([[0, 0, 1, 1, 0, off + len(aenv), off + len(aenv) + 1, 0]], dtype=numpy.int32)
    env2 = numpy.concatenate([env, aenv, [0.0, 1.0]])
    abas[:, 0] += len(atm)
    bas2 = numpy.ascontiguousarray(numpy.vstack([bas, abas, unit]), dtype=numpy.int32)
    atm2 = numpy.ascontiguousarray(numpy.vstack([atm, aatm]), dtype=numpy.int32)
    return atm2, bas2, env2, ao_loc, aux_loc, len(bas), len(abas)


def int3c2e(mol, auxmol):
    """(ij|P), shape (nao, nao, naux) (aux_e2 with aosym='s1',
    pyscf/df/incore.py:34-84)."""
    atm, bas, env, ao_loc, aux_loc, nbas, naux_sh = _df_arrays(mol, auxmol)
    nao, naux = int(ao_loc[-1]), int(aux_loc[-1])
    out = numpy.zeros((nao, nao, naux))
    lib().orc_int3c2e(_p(out), _p(ao_loc), ctypes.c_int(nbas), _p(aux_loc), ctypes.c_int(naux_sh),
                      _p(atm), ctypes.c_int(len(atm)), _p(bas), _p(env))
    return out


def int2c2e(mol, auxmol):
    """(P|Q), shape (naux, naux) (fill_2c2e, pyscf/df/incore.py:110-124)."""
    atm, bas, env, ao_loc, aux_loc, nbas, naux_sh = _df_arrays(mol, auxmol)
    naux = int(aux_loc[-1])
    out = numpy.zeros((naux, naux))
    lib().orc_int2c2e(_p(out), _p(aux_loc), ctypes.c_int(nbas), ctypes.c_int(naux_sh),
                      _p(atm), ctypes.c_int(len(atm)), _p(bas), _p(env))
    return out


def c2s_formula(l):
    nc = (l + 1) * (l + 2) // 2
    out = numpy.zeros((2 * l + 1, nc))
    lib().orc_c2s_formula(ctypes.c_int(l), _p(out))
    return out


def cholesky_eri(mol, auxmol, lindep=1e-7):
    """cderi (naux, nao, nao) = L^-1 (P|ij) with (P|Q) = L L^T; when the
    Cholesky factorisation fails, the eigenvectors of (P|Q) above lindep
    scaled by w^-1/2 (pyscf/df/incore.py:129-214)."""
    import scipy.linalg
    j2c = int2c2e(mol, auxmol)
    j3c = int3c2e(mol, auxmol)
    nao, naux = j3c.shape[0], j3c.shape[2]
    try:
        low = scipy.linalg.cholesky(j2c, lower=True)
        cderi = scipy.linalg.solve_triangular(low, j3c.reshape(-1, naux).T, lower=True)
    except scipy.linalg.LinAlgError:
        w, v = scipy.linalg.eigh(j2c)
        keep = w > lindep
        cderi = (v[:, keep] / numpy.sqrt(w[keep])).T @ j3c.reshape(-1, naux).T
    return cderi.reshape(-1, nao, nao)


def df_get_jk(cderi, dms, hermi=0, with_j=True, with_k=True):
    """DF J and K of real densities (pyscf/df/df_jk.py:278-418):
    J = sum_P B_P tr(B_P D), K = sum_P B_P D B_P."""
    dms = numpy.asarray(dms, dtype=numpy.double)
    shape = dms.shape
    nao = shape[-1]
    dms = dms.reshape(-1, nao, nao)
    vj = vk = None
    if with_j:
        rho = numpy.einsum('pij,nji->np', cderi, dms)
        vj = numpy.einsum('np,pij->nij', rho, cderi).reshape(shape)
    if with_k:
        vk = numpy.array([numpy.einsum('pik,kl,plj->ij', cderi, d, cderi, optimize=True) for d in dms]).reshape(shape)
    return vj, vk
