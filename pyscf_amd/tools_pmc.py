"""Driver for hardware-counter runs: two J/K builds of (H2O)16 def2-SVP (core
guess density), then a 64 MiB device copy whose traffic calibrates the HBM
byte counters.  Run under rocprofv3 --pmc (see profiles/)."""
import sys, os, ctypes
sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import numpy
from pyscf_amd import gto, workloads
from pyscf_amd.scf import _vhf, rhf
import scipy.linalg
mol = gto.M(atom=workloads.water_cluster(16, seed=1), basis='def2-svp')
h = rhf.int1e(mol, 'kin') + rhf.int1e(mol, 'nuc'); s = rhf.int1e(mol, 'ovlp')
e, c = scipy.linalg.eigh(h, s)
dm = 2 * c[:, :mol.nelectron // 2] @ c[:, :mol.nelectron // 2].T
opt = _vhf.VHFOpt(mol, direct_scf_tol=1e-13)
opt.basis.get_jk(dm[None], 1, True, True, 1e-13)
opt.basis.get_jk(dm[None], 1, True, True, 1e-13)
# calibration copy: 64 MiB read + 64 MiB written by k_copy
L = opt.engine.lib; ctx = opt.engine.handle
nb = 64 << 20
a, b = ctypes.c_void_p(), ctypes.c_void_p()
L.mjk_device_alloc(ctx, nb, ctypes.byref(a)); L.mjk_device_alloc(ctx, nb, ctypes.byref(b))
assert L.mjk_device_copy(ctx, b, a, nb) == 0
print('done')
