"""The no-allocation J/K build (mjk_get_jk_device_noalloc) against the
allocating device build (mjk_get_jk_device), both called through the C ABI
on densities already in device memory.

Two no-allocation calls in a row, the second with a smaller input (fewer
AOs, one density instead of two): each must equal the allocating build of
the same input to 1e-10 max-abs on J and K (the FP64 contract; the two runs
share their kernels, so what differs is the order of the floating-point
atomic sums).  Also: the mixed-precision path, and the two refusals that
happen on the host before any kernel runs (task lists built for another
cutoff; work larger than the static scratch).
"""
import ctypes

import numpy
import pytest

from pyscf_amd import gto, workloads
from pyscf_amd.scf import _vhf

pytestmark = pytest.mark.gpu
TOL = 1e-10
SCF_TOL = 1e-13
MJK_EINVAL, MJK_ENOMEM = -1, -2


def _vp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class DeviceArrays:
    """Device copies of the densities plus J and K outputs of the same size."""

    def __init__(self, eng, dms):
        self.eng, self.L, self.ctx = eng, eng.lib, eng.handle
        self.shape = dms.shape
        self.nbytes = dms.nbytes
        self.p = [ctypes.c_void_p() for _ in range(3)]
        for q in self.p:
            eng.check(self.L.mjk_device_alloc(self.ctx, self.nbytes, ctypes.byref(q)), 'mjk_device_alloc')
        eng.check(self.L.mjk_memcpy_h2d(self.ctx, self.p[0], _vp(dms), self.nbytes), 'mjk_memcpy_h2d')

    def read(self):
        self.eng.check(self.L.mjk_synchronize(self.ctx), 'mjk_synchronize')
        vj, vk = numpy.empty(self.shape), numpy.empty(self.shape)
        self.eng.check(self.L.mjk_memcpy_d2h(self.ctx, _vp(vj), self.p[1], self.nbytes), 'mjk_memcpy_d2h')
        self.eng.check(self.L.mjk_memcpy_d2h(self.ctx, _vp(vk), self.p[2], self.nbytes), 'mjk_memcpy_d2h')
        return vj, vk

    def free(self):
        for q in self.p:
            self.L.mjk_device_free(self.ctx, q)


def build(basis, dev, hermi, noalloc, tol=SCF_TOL, fp32_tol=0.0):
    L, ctx = basis.engine.lib, basis.engine.handle
    n_dm = dev.shape[0]
    d, j, k = dev.p
    if noalloc:
        return L.mjk_get_jk_device_noalloc(ctx, basis.handle, d, n_dm, hermi, 1, 1, tol, fp32_tol, j, k)
    return L.mjk_get_jk_device(ctx, basis.handle, d, n_dm, hermi, 1, 1, tol, fp32_tol, j, k, None)


def densities(nao, n_dm, seed, symmetric):
    rng = numpy.random.default_rng(seed)
    out = []
    for _ in range(n_dm):
        c = rng.standard_normal((nao, max(2, nao // 4))) / 4
        d = c @ c.T
        if not symmetric:
            d = d + 0.1 * rng.standard_normal((nao, nao))
        out.append(d)
    return numpy.ascontiguousarray(numpy.array(out))


def reference_and_noalloc(basis, dms, hermi, fp32_tol=0.0):
    eng = basis.engine
    ref = DeviceArrays(eng, dms)
    out = DeviceArrays(eng, dms)
    try:
        eng.check(build(basis, ref, hermi, False, fp32_tol=fp32_tol), 'mjk_get_jk_device')
        r = ref.read()
        eng.check(build(basis, out, hermi, True, fp32_tol=fp32_tol), 'mjk_get_jk_device_noalloc')
        o = out.read()
    finally:
        ref.free()
        out.free()
    return r, o


def test_noalloc_two_calls_second_smaller():
    eng = _vhf.get_engine()
    big = gto.M(atom=workloads.water_cluster(8, seed=1), basis='def2-svp')
    small = gto.M(atom=workloads.h2o(), basis='cc-pvdz')
    b_big = eng.upload(big._atm, big._bas, big._env, SCF_TOL)
    b_small = eng.upload(small._atm, small._bas, small._env, SCF_TOL)
    d_big = densities(big.nao_nr(), 2, 11, symmetric=False)
    d_small = densities(small.nao_nr(), 1, 12, symmetric=True)
    # references first, then the two no-allocation calls back to back
    ref_big, ref_small = DeviceArrays(eng, d_big), DeviceArrays(eng, d_small)
    out_big, out_small = DeviceArrays(eng, d_big), DeviceArrays(eng, d_small)
    try:
        eng.check(build(b_big, ref_big, 0, False), 'mjk_get_jk_device')
        eng.check(build(b_small, ref_small, 1, False), 'mjk_get_jk_device')
        eng.check(build(b_big, out_big, 0, True), 'mjk_get_jk_device_noalloc (first)')
        eng.check(build(b_small, out_small, 1, True), 'mjk_get_jk_device_noalloc (second, smaller)')
        for name, ref, out in (('first', ref_big, out_big), ('second', ref_small, out_small)):
            (j0, k0), (j1, k1) = ref.read(), out.read()
            dj, dk = abs(j1 - j0).max(), abs(k1 - k0).max()
            print('noalloc %s call: max|dJ|=%.2e max|dK|=%.2e' % (name, dj, dk), flush=True)
            assert numpy.isfinite(j1).all() and numpy.isfinite(k1).all()
            assert abs(j0).max() > 1e-3 and abs(k0).max() > 1e-3
            assert dj <= TOL and dk <= TOL, name
    finally:
        for a in (ref_big, ref_small, out_big, out_small):
            a.free()


def test_noalloc_mixed_precision_and_unscreened():
    eng = _vhf.get_engine()
    mol = gto.M(atom=workloads.water_cluster(4, seed=2), basis='def2-svp')
    b = eng.upload(mol._atm, mol._bas, mol._env, SCF_TOL)
    dms = densities(mol.nao_nr(), 1, 13, symmetric=True)
    (j0, k0), (j1, k1) = reference_and_noalloc(b, dms, 1, fp32_tol=1e-7)
    assert abs(j1 - j0).max() <= TOL and abs(k1 - k0).max() <= TOL
    # an unscreened basis (cutoff 0) builds unscreened
    b0 = eng.upload(mol._atm, mol._bas, mol._env, 0.0)
    dev0, dev1 = DeviceArrays(eng, dms), DeviceArrays(eng, dms)
    try:
        eng.check(build(b0, dev0, 1, False, tol=0.0), 'mjk_get_jk_device')
        eng.check(build(b0, dev1, 1, True, tol=0.0), 'mjk_get_jk_device_noalloc')
        (j0, k0), (j1, k1) = dev0.read(), dev1.read()
        assert abs(j1 - j0).max() <= TOL and abs(k1 - k0).max() <= TOL
    finally:
        dev0.free()
        dev1.free()


def test_noalloc_refusals():
    eng = _vhf.get_engine()
    mol = gto.M(atom=workloads.h2o(), basis='cc-pvdz')
    b = eng.upload(mol._atm, mol._bas, mol._env, SCF_TOL)
    dms = densities(mol.nao_nr(), 1, 14, symmetric=True)
    dev = DeviceArrays(eng, dms)
    try:
        # task lists exist for 1e-13 only: another cutoff is refused, nothing built
        assert build(b, dev, 1, True, tol=1e-10) == MJK_EINVAL
        assert 'task lists' in eng.error()
    finally:
        dev.free()
    # 6000 densities with hermi=0 need about 4e8 bytes of work buffers
    many = numpy.zeros((6000,) + dms.shape[1:])
    dev = DeviceArrays(eng, many)
    try:
        assert build(b, dev, 0, True) == MJK_ENOMEM
        assert 'scratch' in eng.error()
    finally:
        dev.free()
