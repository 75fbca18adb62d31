"""One-electron integrals on the GPU (mjk_int1e_device) against the host
code (mjk_int1e, pinned against the oracle and the reference's energies in
tests/test_int1e.py): overlap, kinetic and nuclear-attraction matrices for
s to f shells, generally contracted shells and many atoms, host and device
outputs."""
import ctypes

import numpy
import pytest

from pyscf_amd import gto, workloads
from pyscf_amd.scf import _vhf, rhf

pytestmark = pytest.mark.gpu

MOLS = {
    'h2o_vdz': lambda: gto.M(atom=workloads.h2o(), basis='cc-pvdz'),
    'chain_tzvp': lambda: gto.M(atom=workloads.random_chain([9, 3], nmin=5, nmax=5), basis='def2-tzvp'),
    'water8_svp': lambda: gto.M(atom=workloads.water_cluster(8, seed=2), basis='def2-svp'),
}


@pytest.mark.parametrize('name', sorted(MOLS))
def test_int1e_device_equals_host(name):
    mol = MOLS[name]()
    dev = rhf.int1e_device(mol)
    for which in ('ovlp', 'kin', 'nuc'):
        host = rhf.int1e(mol, which)
        err = abs(dev[which] - host).max()
        print(name, which, mol.nao_nr(), 'max|d|=%.2e (max %.2e)' % (err, abs(host).max()), flush=True)
        assert err <= 1e-12 * max(1.0, abs(host).max())
        assert abs(dev[which] - dev[which].T).max() <= 1e-13 * max(1.0, abs(host).max())


def test_int1e_device_outputs():
    mol = MOLS['h2o_vdz']()
    eng = _vhf.get_engine()
    L = eng.lib
    nao = mol.nao_nr()
    ptrs = []
    for _ in range(3):
        ptr = ctypes.c_void_p()
        eng.check(L.mjk_device_alloc(eng.handle, 8 * nao * nao, ctypes.byref(ptr)), 'alloc')
        ptrs.append(ptr)
    a = [numpy.ascontiguousarray(x, dtype=t) for x, t in ((mol._atm, numpy.int32), (mol._bas, numpy.int32),
                                                          (mol._env, numpy.double))]
    p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    eng.check(L.mjk_int1e_device(eng.handle, p(a[0]), len(a[0]), p(a[1]), len(a[1]), p(a[2]), ptrs[0], ptrs[1],
                                 ptrs[2], 1), 'mjk_int1e_device')
    eng.check(L.mjk_synchronize(eng.handle), 'sync')
    for ptr, which in zip(ptrs, ('ovlp', 'kin', 'nuc')):
        out = numpy.empty((nao, nao))
        eng.check(L.mjk_memcpy_d2h(eng.handle, p(out), ptr, out.nbytes), 'd2h')
        assert abs(out - rhf.int1e(mol, which)).max() <= 1e-12
        L.mjk_device_free(eng.handle, ptr)
