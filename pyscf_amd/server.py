"""Continuous-batching request server for J/K builds and whole RHF solves,
over the library's C server (mjk_server_create / mjk_submit / mjk_poll /
mjk_wait_any / mjk_result).

* A request is ``('get_jk', mol, dm, hermi)`` or ``('rhf', mol)``.
* The C server owns ``nslots`` worker contexts (one stream each, default
  priority); requests running together overlap on the device.  An RHF runs
  the device-resident SCF: density, Fock matrix and DIIS history stay in
  HBM for the request's life.
* Admission in arrival order on the device bytes the contexts track for the
  running requests (each at least at its estimate until its buffers exist),
  against ``hbm_fraction`` of the device memory; a request whose estimate
  alone exceeds the budget fails with MemoryError.
* Each finished request records its submission, start and finish times on
  the server's clock (seconds), so the caller computes p50/p99 latency and
  throughput.
"""

import ctypes
import threading
from concurrent.futures import Future

import numpy

from pyscf_amd import lib as _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Request:
    def __init__(self, kind, mol, dm=None, hermi=1, tag=None):
        if kind not in ('get_jk', 'rhf'):
            raise ValueError(kind)
        self.kind = kind
        self.mol = mol
        self.dm = dm
        self.hermi = hermi
        self.tag = tag
        self.ticket = None
        self.t_arrival = self.t_start = self.t_done = None
        self.info = None
        self.future = Future()


class JKServer:
    def __init__(self, nslots=8, hbm_fraction=0.8, device=0, direct_scf_tol=1e-13, fp32_tol=0.0,
                 conv_tol=1e-9, max_cycle=100):
        self.lib = L = _lib.load_library()
        free, total = ctypes.c_size_t(), ctypes.c_size_t()
        if L.mjk_mem_info(device, ctypes.byref(free), ctypes.byref(total)) != 0:
            raise RuntimeError('no HIP device %d' % device)
        self.budget = max(1, int(hbm_fraction * total.value))
        self.direct_scf_tol = direct_scf_tol
        self.fp32_tol = fp32_tol
        self.conv_tol = conv_tol
        self.max_cycle = max_cycle
        h = ctypes.c_void_p()
        rc = L.mjk_server_create(device, nslots, self.budget, ctypes.byref(h))
        if rc != 0:
            raise RuntimeError('mjk_server_create failed (code %d)' % rc)
        self.handle = h
        self.lock = threading.Lock()
        self.live = {}          # ticket -> Request
        self.finished = []
        self.closed = False
        self.collector = threading.Thread(target=self._collect, daemon=True)
        self.collector.start()

    def hbm_held(self):
        return self.lib.mjk_server_hbm_held(self.handle)

    def submit(self, req):
        L = self.lib
        mol = req.mol
        keep = [numpy.ascontiguousarray(mol._atm, dtype=numpy.int32),
                numpy.ascontiguousarray(mol._bas, dtype=numpy.int32),
                numpy.ascontiguousarray(mol._env, dtype=numpy.double)]
        q = _lib.mjk_request()
        q.kind = 0 if req.kind == 'get_jk' else 1
        q.atm, q.natm = keep[0].ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(keep[0])
        q.bas, q.nbas = keep[1].ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(keep[1])
        q.env, q.env_size = keep[2].ctypes.data_as(ctypes.POINTER(ctypes.c_double)), keep[2].size
        q.direct_scf_tol = self.direct_scf_tol
        q.fp32_tol = self.fp32_tol
        nao = mol.nao_nr()
        if req.kind == 'get_jk':
            dm = numpy.ascontiguousarray(req.dm, dtype=numpy.double)
            if dm.size % (nao * nao):
                raise ValueError('density shape %s does not fit %d AOs' % (dm.shape, nao))
            keep.append(dm)
            q.dm = dm.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            q.n_dm = dm.size // (nao * nao)
            q.hermi, q.with_j, q.with_k = int(req.hermi), 1, 1
        else:
            q.nelectron = mol.nelectron
            q.max_cycle = self.max_cycle
            q.conv_tol = self.conv_tol
        t = ctypes.c_longlong()
        with self.lock:
            rc = L.mjk_submit(self.handle, ctypes.byref(q), ctypes.byref(t))
            if rc == 0:
                req.ticket = t.value
                self.live[t.value] = req
        if rc != 0:
            msg = L.mjk_server_error(self.handle).decode()
            req.future.set_exception(MemoryError(msg) if rc == -2 else ValueError(msg))
        return req.future

    def _collect(self):
        L = self.lib
        t = ctypes.c_longlong()
        while True:
            with self.lock:
                if self.closed and not self.live:
                    return
            L.mjk_wait_any(self.handle, 0.05, ctypes.byref(t))
            if t.value == 0:
                continue
            with self.lock:
                req = self.live.pop(t.value)
            info = _lib.mjk_request_info()
            L.mjk_poll(self.handle, t.value, ctypes.byref(info))
            req.info = info
            req.t_arrival, req.t_start, req.t_done = info.t_submit, info.t_start, info.t_done
            n, ndm = info.nao, max(info.n_dm, 1)
            if req.kind == 'get_jk':
                shape = (n, n) if numpy.asarray(req.dm).ndim == 2 else (ndm, n, n)
                vj, vk = numpy.empty(shape), numpy.empty(shape)
                rc = L.mjk_result(self.handle, t.value, _ptr(vj), _ptr(vk), None, None)
                out = (vj, vk)
            else:
                dm, mo_energy = numpy.empty((n, n)), numpy.empty(n)
                rc = L.mjk_result(self.handle, t.value, None, None, _ptr(dm), _ptr(mo_energy))
                req.dm_out, req.mo_energy = dm, mo_energy
                out = (info.e_tot, bool(info.converged))
            with self.lock:
                self.finished.append(req)
            if rc != 0:
                req.future.set_exception(RuntimeError(L.mjk_server_error(self.handle).decode()))
            else:
                req.future.set_result(out)

    def close(self, wait=True):
        with self.lock:
            self.closed = True
        if wait:
            self.collector.join()
            self.lib.mjk_server_destroy(self.handle)
            self.handle = None
