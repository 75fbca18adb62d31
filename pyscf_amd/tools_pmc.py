"""Driver for hardware-counter runs of the (H2O)16 def2-SVP J/K build.

    rocprofv3 --pmc FETCH_SIZE -f csv -d OUT -o pmc -- python3 pyscf_amd/tools_pmc.py DM.npy

Two J/K builds of the density saved by ``bench.py --dm-out DM.npy`` (the
converged RHF density the benchmark times), then a 64 MiB device copy
(kernel ``k_copy``) whose byte counts calibrate the HBM counters: it moves
exactly 64 MiB in each direction.  Counter values per launch are the sums
over the two builds halved.
"""
import ctypes
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import numpy

from pyscf_amd import gto, workloads
from pyscf_amd.scf import _vhf


def main(path):
    mol = gto.M(atom=workloads.water_cluster(16, seed=1), basis='def2-svp')
    dm = numpy.ascontiguousarray(numpy.load(path, allow_pickle=False))
    assert dm.shape == (mol.nao_nr(), mol.nao_nr())
    opt = _vhf.VHFOpt(mol, direct_scf_tol=1e-13)
    for _ in range(2):
        opt.basis.get_jk(dm[None], 1, True, True, 1e-13)
    lib, ctx = opt.engine.lib, opt.engine.handle
    nb = 64 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    opt.engine.check(lib.mjk_device_alloc(ctx, nb, ctypes.byref(a)), 'alloc')
    opt.engine.check(lib.mjk_device_alloc(ctx, nb, ctypes.byref(b)), 'alloc')
    opt.engine.check(lib.mjk_device_copy(ctx, b, a, nb), 'copy')
    opt.engine.check(lib.mjk_synchronize(ctx), 'sync')
    print('done', flush=True)


if __name__ == '__main__':
    main(sys.argv[1])
