"""Bindings of the HIP J/K engine, mirroring pyscf/scf/_vhf.py.

* ``Engine``   -- one device context (stream, HBM budget, Rys table).
* ``VHFOpt``   -- screening context of one molecule; counterpart of
                  ``_VHFOpt`` (pyscf/scf/_vhf.py:151-264): Schwarz bounds
                  (q_cond) are computed once on the GPU and stay resident
                  with the sorted shell pairs; dm_cond is built per call on
                  the device from the density.
* ``direct``   -- counterpart of ``_vhf.direct`` (pyscf/scf/_vhf.py:405-464):
                  J with 'ji->s2kl', K with 'li->s2kj' (hermi=1) or
                  'li->s1kj', then hermi_triu on the outputs.

There is no CPU fallback: without the built library or a HIP device every
call raises.
"""

import collections
import ctypes
import hashlib
import threading

import numpy

from pyscf_amd import lib as _lib

PTR_RANGE_OMEGA = 8       # env slot of the two-electron operator (pyscf/gto/mole.py:88)

_engines = {}
_engine_lock = threading.Lock()


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Engine:
    """Device context of the J/K engine (one per HIP device and process)."""

    def __init__(self, device=0, hbm_budget=0):
        self.lib = _lib.load_library()
        self.device = device
        h = ctypes.c_void_p()
        rc = self.lib.mjk_ctx_create(ctypes.byref(h), device, hbm_budget)
        if rc != 0:
            raise RuntimeError('cannot create the J/K engine on HIP device %d (code %d); a GPU is required'
                               % (device, rc))
        self.handle = h
        self._lock = threading.RLock()
        self._basis_cache = collections.OrderedDict()

    def error(self):
        return self.lib.mjk_last_error(self.handle).decode()

    def check(self, rc, what):
        if rc != 0:
            raise RuntimeError('%s failed (code %d): %s' % (what, rc, self.error()))

    def hbm_used(self):
        return self.lib.mjk_ctx_hbm_used(self.handle)

    def upload(self, atm, bas, env, cutoff):
        atm = numpy.ascontiguousarray(atm, dtype=numpy.int32)
        bas = numpy.ascontiguousarray(bas, dtype=numpy.int32)
        env = numpy.ascontiguousarray(env, dtype=numpy.double)
        h = ctypes.c_void_p()
        rc = self.lib.mjk_basis_upload(self.handle, _p(atm), len(atm), _p(bas), len(bas), _p(env),
                                       float(cutoff), ctypes.byref(h))
        self.check(rc, 'mjk_basis_upload')
        return Basis(self, h, [int(self.lib.mjk_basis_nao(h, 0))], nbas=[len(bas)])

    def upload_batch(self, mols, cutoff):
        arrs = []
        for m in mols:
            arrs.append((numpy.ascontiguousarray(m._atm, dtype=numpy.int32),
                         numpy.ascontiguousarray(m._bas, dtype=numpy.int32),
                         numpy.ascontiguousarray(m._env, dtype=numpy.double)))
        n = len(arrs)
        natm = numpy.array([len(a[0]) for a in arrs], dtype=numpy.int32)
        nbas = numpy.array([len(a[1]) for a in arrs], dtype=numpy.int32)
        patm = (ctypes.c_void_p * n)(*[a[0].ctypes.data for a in arrs])
        pbas = (ctypes.c_void_p * n)(*[a[1].ctypes.data for a in arrs])
        penv = (ctypes.c_void_p * n)(*[a[2].ctypes.data for a in arrs])
        h = ctypes.c_void_p()
        rc = self.lib.mjk_basis_upload_batch(self.handle, n, _p(natm), _p(nbas), patm, pbas, penv,
                                             float(cutoff), ctypes.byref(h))
        self.check(rc, 'mjk_basis_upload_batch')
        naos = [int(self.lib.mjk_basis_nao(h, i)) for i in range(n)]
        return Basis(self, h, naos, nbas=list(nbas))

    def cached_basis(self, atm, bas, env):
        """Unscreened basis handle for the module-level get_jk, cached by
        content (a small LRU)."""
        key = hashlib.sha1(numpy.ascontiguousarray(atm, dtype=numpy.int32).tobytes()
                           + numpy.ascontiguousarray(bas, dtype=numpy.int32).tobytes()
                           + numpy.ascontiguousarray(env, dtype=numpy.double).tobytes()).hexdigest()
        with self._lock:
            b = self._basis_cache.get(key)
            if b is None:
                b = self.upload(atm, bas, env, 0.0)
                self._basis_cache[key] = b
                while len(self._basis_cache) > 16:
                    self._basis_cache.popitem(last=False)
            else:
                self._basis_cache.move_to_end(key)
            return b


def get_engine(device=0):
    with _engine_lock:
        e = _engines.get(device)
        if e is None:
            e = Engine(device)
            _engines[device] = e
        return e


class Basis:
    """Device-resident shell pairs and work buffers of one molecule or batch."""

    def __init__(self, engine, handle, naos, nbas):
        self.engine = engine
        self.handle = handle
        self.naos = naos
        self.nbas = nbas
        self.slot = int(engine.lib.mjk_basis_slot_size(handle))
        self.last_stats = None

    def __del__(self):
        try:
            if self.handle:
                self.engine.lib.mjk_basis_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def set_active(self, active):
        """Batched builds: molecules with active[m] false are skipped at block
        level by the following builds (their J and K come back zero)."""
        a = numpy.ascontiguousarray(numpy.asarray(active, dtype=bool).astype(numpy.int32))
        if a.size != len(self.naos):
            raise ValueError('need one flag per molecule (%d), got %d' % (len(self.naos), a.size))
        self.engine.check(self.engine.lib.mjk_basis_set_active(self.engine.handle, self.handle, _p(a)),
                          'mjk_basis_set_active')

    def q_cond(self, mol=0):
        n = self.nbas[mol]
        out = numpy.zeros((n, n))
        self.engine.check(self.engine.lib.mjk_basis_q_cond(self.engine.handle, self.handle, mol, _p(out)),
                          'mjk_basis_q_cond')
        return out

    def get_jk(self, dms, hermi, with_j, with_k, direct_scf_tol, fp32_tol=0.0, stats=False):
        """dms: float64 array (n_dm, slot) or (n_dm, nao, nao) for one molecule."""
        dms = numpy.ascontiguousarray(dms, dtype=numpy.double)
        n_dm = dms.shape[0]
        if dms.size != n_dm * self.slot:
            raise ValueError('density shape %s does not match the basis (%d elements per slot)'
                             % (dms.shape, self.slot))
        # zero-filled: with inactive molecules the library writes only the active blocks
        vj = numpy.zeros_like(dms) if with_j else None
        vk = numpy.zeros_like(dms) if with_k else None
        st = _lib.mjk_stats()
        rc = self.engine.lib.mjk_get_jk(self.engine.handle, self.handle, _p(dms), n_dm, int(hermi),
                                        int(bool(with_j)), int(bool(with_k)), float(direct_scf_tol),
                                        float(fp32_tol), _p(vj), _p(vk),
                                        ctypes.byref(st) if stats else None)
        self.engine.check(rc, 'mjk_get_jk')
        if stats:
            self.last_stats = st.as_dict()
        return vj, vk


    def get_jk_resident(self, n_dm, hermi, with_j, with_k, direct_scf_tol, fp32_tol=0.0, stats=False):
        """The build of the densities the last get_jk uploaded, J and K left
        on the device (no host copies); returns the statistics with stats."""
        st = _lib.mjk_stats()
        rc = self.engine.lib.mjk_get_jk_resident(self.engine.handle, self.handle, int(n_dm), int(hermi),
                                                 int(bool(with_j)), int(bool(with_k)), float(direct_scf_tol),
                                                 float(fp32_tol), ctypes.byref(st) if stats else None)
        self.engine.check(rc, 'mjk_get_jk_resident')
        if stats:
            self.last_stats = st.as_dict()
            return self.last_stats
        return None

    def fetch_jk(self, n_dm, with_j=True, with_k=True):
        """Resident J and K of the last build, as (n_dm, slot) arrays."""
        vj = numpy.empty((n_dm, self.slot)) if with_j else None
        vk = numpy.empty((n_dm, self.slot)) if with_k else None
        self.engine.check(self.engine.lib.mjk_basis_fetch_jk(self.engine.handle, self.handle, int(n_dm),
                                                             _p(vj), _p(vk)), 'mjk_basis_fetch_jk')
        return vj, vk

    def set_class_filter(self, cls):
        """Following builds run quartet class cls only (-1: all classes)."""
        self.engine.check(self.engine.lib.mjk_basis_set_class_filter(self.engine.handle, self.handle, int(cls)),
                          'mjk_basis_set_class_filter')

    def classes(self, n_dm=1, hermi=1):
        """[(index, (la, lb, lc, ld), (ca, cb, cc, cd), nblocks, kernel)] of the
        class list; ctr = contraction counts of the shells (2: a generally
        contracted s shell kept whole); kernel 0 tile, 1 row, 2 cooperative,
        3 merged row kernel, 4 row kernel with the B rows of K' in memory, -1
        the class does not run in such a build."""
        L = self.engine.lib
        out = []
        l4 = numpy.zeros(4, dtype=numpy.int32)
        c4 = numpy.zeros(4, dtype=numpy.int32)
        nb, kern = ctypes.c_longlong(), ctypes.c_int()
        for k in range(L.mjk_basis_nclasses(self.handle)):
            self.engine.check(L.mjk_basis_class(self.engine.handle, self.handle, k, n_dm, hermi, _p(l4), _p(c4),
                                                ctypes.byref(nb), ctypes.byref(kern)), 'mjk_basis_class')
            out.append((k, tuple(int(x) for x in l4), tuple(int(x) for x in c4), int(nb.value), int(kern.value)))
        return out


class VHFOpt:
    """Screening context (counterpart of _VHFOpt, pyscf/scf/_vhf.py:151-264).

    The Schwarz bound q_cond and the shell-pair data are built once on the
    device; ``direct_scf_tol`` is the prescreen cutoff.  The predicate
    follows the requested matrices: with J and K it is CVHFnrs8_prescreen
    (pyscf/lib/vhf/optimizer.c:90-117), with J only CVHFnrs8_vj_prescreen,
    with K only CVHFnrs8_vk_prescreen (:119-167), as SCF.get_jk swaps them
    (pyscf/scf/hf.py:2069-2077)."""

    def __init__(self, mol, intor=None, prescreen='CVHFnrs8_prescreen', qcondname='CVHFnr_int2e_q_cond',
                 dmcondname='CVHFnr_dm_cond', direct_scf_tol=1e-14, engine=None):
        if prescreen not in ('CVHFnrs8_prescreen', 'CVHFnrs8_vj_prescreen', 'CVHFnrs8_vk_prescreen'):
            raise NotImplementedError('prescreen %s' % prescreen)
        self.mol = mol
        self.prescreen = prescreen
        self.direct_scf_tol = direct_scf_tol
        self.engine = engine or get_engine()
        self.omega = float(mol._env[PTR_RANGE_OMEGA])   # the bounds are of this operator
        self.basis = self.engine.upload(mol._atm, mol._bas, mol._env, direct_scf_tol)

    @property
    def q_cond(self):
        return self.basis.q_cond(0)


def direct(dms, atm, bas, env, vhfopt=None, hermi=0, cart=False, with_j=True, with_k=True, out=None,
           fp32_tol=0.0):
    """J and K for a stack of real densities (pyscf/scf/_vhf.py:405-464)."""
    dms = numpy.asarray(dms, order='C', dtype=numpy.double)
    dms_shape = dms.shape
    nao = dms_shape[-1]
    dms = dms.reshape(-1, nao, nao)
    if cart and int(numpy.asarray(bas)[:, 1].max()) > 1:
        raise NotImplementedError('Cartesian d/f functions are not supported by the J/K engine')
    if not (with_j or with_k):
        return None, None
    if vhfopt is None:
        basis = get_engine().cached_basis(atm, bas, env)
        tol = 0.0
    else:
        if float(env[PTR_RANGE_OMEGA]) != vhfopt.omega:
            raise ValueError('vhfopt was built for omega=%g but env[PTR_RANGE_OMEGA]=%g'
                             % (vhfopt.omega, float(env[PTR_RANGE_OMEGA])))
        basis = vhfopt.basis
        tol = vhfopt.direct_scf_tol
    if basis.naos[0] != nao:
        raise ValueError('dm has nao=%d but the basis has %d AOs' % (nao, basis.naos[0]))
    vj, vk = basis.get_jk(dms, hermi, with_j, with_k, tol, fp32_tol)
    if with_j:
        vj = vj.reshape(dms_shape)
    if with_k:
        vk = vk.reshape(dms_shape)
    return vj, vk
