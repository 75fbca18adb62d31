"""Build the HIP library ``pyscf_amd/lib/libmjk.so`` for gfx950 in-tree.

    python -m pyscf_amd.build

Each .hip translation unit is compiled in parallel, then linked.  Flags:
gfx950 only, -O3, packed FP32 instructions disabled (the compiler would
otherwise emit v_pk_* FP32 ops), FP64 atomics lowered to the hardware
instruction (-munsafe-fp-atomics).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT = os.path.join(HERE, 'lib', 'libmjk.so')
BUILD = os.path.join(HERE, 'csrc', '_build')
SOURCES = ['jk_kernels.hip', 'mjk_api.hip']
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-munsafe-fp-atomics', '-Wno-unused-result']


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + '.o')
    cmd = [HIPCC] + FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s' % (src, r.stderr[-4000:]))
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    with ThreadPoolExecutor(len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + '.tmp'
    cmd = [HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n%s' % r.stderr[-4000:])
    os.replace(tmp, OUT)
    if verbose:
        print('built', OUT)
    return OUT


if __name__ == '__main__':
    build()
    sys.exit(0)
