"""Continuous-batching request loop for J/K builds and whole RHF solves.

* A request is ``('get_jk', mol, dm, hermi)`` or ``('rhf', mol)``.
* The server owns ``nslots`` slots; each slot is a device context with its
  own HIP stream (default priority) and runs one request at a time on a
  worker thread.  The C calls release the GIL, so requests that arrive
  while others are on the GPU run concurrently with them: kernels of
  different requests overlap on the device.
* Admission: a request starts when a slot is free AND its HBM estimate fits
  under ``hbm_fraction`` of the device memory together with the running
  requests; otherwise it waits in arrival order.  Shell pairs, Schwarz bounds
  and densities stay resident in the slot's context for the request's life.
* Each finished request records arrival, start and finish times, so the
  caller computes p50/p99 latency and throughput.
"""

import ctypes
import queue
import threading
import time
from concurrent.futures import Future

import numpy

from pyscf_amd import lib as _lib
from pyscf_amd.scf import _vhf, hf, rhf


def hbm_estimate(mol):
    """Upper bound of the device bytes a request holds (pairs, tiles, arenas)."""
    bas = mol._bas
    l = bas[:, 1]
    nctr = bas[:, 3]
    nprim = bas[:, 2]
    # segmented shells
    seg_np = numpy.repeat(nprim, nctr)
    seg_l = numpy.repeat(l, nctr)
    nsh = len(seg_np)
    nprim_pairs = (numpy.sum(seg_np) ** 2 + numpy.sum(seg_np ** 2)) // 2
    npairs = nsh * (nsh + 1) // 2
    ncart = int(numpy.sum((seg_l + 1) * (seg_l + 2) // 2))
    nao = int(numpy.sum(2 * seg_l + 1))
    pair_bytes = 48 * nprim_pairs + 120 * npairs
    tiles = 16 * 55 * 16 * (npairs // 8 + 1) * 2
    arenas = 8 * (6 * ncart * ncart + 3 * nao * nao + nsh * nsh)
    return int(pair_bytes + tiles + arenas + (1 << 20))


class Request:
    def __init__(self, kind, mol, dm=None, hermi=1, tag=None):
        if kind not in ('get_jk', 'rhf'):
            raise ValueError(kind)
        self.kind = kind
        self.mol = mol
        self.dm = dm
        self.hermi = hermi
        self.tag = tag
        self.t_arrival = None
        self.t_start = None
        self.t_done = None
        self.hbm = hbm_estimate(mol)
        self.future = Future()


class JKServer:
    def __init__(self, nslots=8, hbm_fraction=0.8, device=0, direct_scf_tol=1e-13, fp32_tol=0.0,
                 conv_tol=1e-9):
        self.lib = _lib.load_library()
        free, total = ctypes.c_size_t(), ctypes.c_size_t()
        if self.lib.mjk_mem_info(device, ctypes.byref(free), ctypes.byref(total)) != 0:
            raise RuntimeError('no HIP device %d' % device)
        self.budget = int(hbm_fraction * total.value)
        self.direct_scf_tol = direct_scf_tol
        self.fp32_tol = fp32_tol
        self.conv_tol = conv_tol
        # admission enforces the budget; each slot context may use what it is admitted with
        self.engines = [_vhf.Engine(device) for _ in range(nslots)]
        self.free_slots = list(range(nslots))
        self.pending = []
        self.in_use = 0
        self.lock = threading.Condition()
        self.closed = False
        self.finished = []
        self.dispatcher = threading.Thread(target=self._dispatch, daemon=True)
        self.dispatcher.start()

    def submit(self, req):
        if req.hbm > self.budget:
            req.future.set_exception(MemoryError('request needs %d B, budget %d B' % (req.hbm, self.budget)))
            return req.future
        with self.lock:
            req.t_arrival = time.perf_counter()
            self.pending.append(req)
            self.lock.notify_all()
        return req.future

    def _dispatch(self):
        while True:
            with self.lock:
                while not self.closed and not self._admissible():
                    self.lock.wait()
                if self.closed and not self.pending:
                    return
                if not self._admissible():
                    continue
                req = self.pending.pop(0)
                slot = self.free_slots.pop()
                self.in_use += req.hbm
            threading.Thread(target=self._run, args=(req, slot), daemon=True).start()

    def _admissible(self):
        return bool(self.pending) and bool(self.free_slots) and self.in_use + self.pending[0].hbm <= self.budget

    def _run(self, req, slot):
        eng = self.engines[slot]
        req.t_start = time.perf_counter()
        try:
            if req.kind == 'get_jk':
                opt = _vhf.VHFOpt(req.mol, direct_scf_tol=self.direct_scf_tol, engine=eng)
                out = hf.get_jk(req.mol, req.dm, req.hermi, opt, fp32_tol=self.fp32_tol)
                del opt
            else:
                mf = rhf.RHF(req.mol, direct_scf_tol=self.direct_scf_tol, fp32_tol=self.fp32_tol)
                mf._opt = {None: _vhf.VHFOpt(req.mol, direct_scf_tol=self.direct_scf_tol, engine=eng)}
                mf.conv_tol = self.conv_tol
                e = mf.kernel()
                out = (e, mf.converged)
                mf._opt = None
            req.t_done = time.perf_counter()
            req.future.set_result(out)
        except Exception as exc:      # the request fails, the server keeps serving
            req.t_done = time.perf_counter()
            req.future.set_exception(exc)
        finally:
            with self.lock:
                self.in_use -= req.hbm
                self.free_slots.append(slot)
                self.finished.append(req)
                self.lock.notify_all()

    def close(self, wait=True):
        with self.lock:
            self.closed = True
            self.lock.notify_all()
        if wait:
            self.dispatcher.join()
            while True:
                with self.lock:
                    if len(self.free_slots) == len(self.engines):
                        break
                time.sleep(0.01)
