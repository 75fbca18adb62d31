"""Build the HIP library ``pyscf_amd/lib/libmjk.so`` for gfx950 in-tree.

    python -m pyscf_amd.build [--force]

The quartet-class kernels are instantiated one translation unit per class
(generated into csrc/_build/) so that the slow template instantiations
compile in parallel; the dispatch table maps (la lb | lc ld) to them.
Flags: gfx950 only, -O3, packed FP32 instructions disabled, FP64 atomics
lowered to the hardware instruction (-munsafe-fp-atomics).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT = os.path.join(HERE, 'lib', 'libmjk.so')
BUILD = os.path.join(CSRC, '_build')
INCLUDE = os.path.join(os.path.dirname(HERE), 'include')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-munsafe-fp-atomics', '-Wno-unused-result', '-I' + CSRC, '-I' + INCLUDE]
# dense algebra of the device-resident SCF
LIBS = ['-L/opt/rocm/lib', '-lrocsolver', '-lrocblas', '-Wl,-rpath,/opt/rocm/lib']
STATIC_SOURCES = ['jk_misc.hip', 'mjk_api.hip', 'mjk_scf.hip', 'mjk_serve.hip', 'mjk_df.hip', 'mjk_int1e.hip']
LMAX = 3


def quartet_classes(lmax=LMAX):
    pairs = [(la, lb) for la in range(lmax + 1) for lb in range(la + 1)]
    out = []
    for x in range(len(pairs)):
        for y in range(x + 1):
            out.append(pairs[x] + pairs[y])
    return out


def _name(c):
    return 'jk_%d%d%d%d' % c


def ncart(l):
    return (l + 1) * (l + 2) // 2


def jk_split(la, lb, lc, ld):
    """Slices of a quartet class (same rule as jk_split in jk_device.h)."""
    na, nb, nc, nd = ncart(la), ncart(lb), ncart(lc), ncart(ld)
    nf = na * nb * nc * nd
    if nf <= 81:
        return 1
    if (la, lb, lc, ld) == (2, 1, 1, 1):     # 81 per lane: one wave per SIMD either way
        return 2
    for s in range(2, na + 1):
        if na % s == 0 and nf // s <= 54:
            return s
    if nf // (na * nb) <= 54:
        return na * nb
    return na * nb * nc


SLICES_PER_UNIT = 3


def nfunc(c):
    return ncart(c[0]) * ncart(c[1]) * ncart(c[2]) * ncart(c[3])


def use_row(c):
    """Classes that get the row kernel (one block per bra-pair task, K' rows
    accumulated in LDS): at most 120 Cartesian components per quartet, or at
    most 180 with at most 9 ket components (the J'[c,d] updates stay per
    quartet), cut into at most 6 slices.  Measured on MI355X ((H2O)16 def2-SVP,
    ms per launch): (ds|ds) row 0.18 / cooperative 0.97, (dd|ss) 0.12 / 0.64,
    (dp|pp) 0.84 / 1.11; (dd|ds) 0.29 / 0.16 and, in def2-TZVP, (fs|dp)
    2.41 / 2.04 stay cooperative."""
    nf = nfunc(c)
    nket = ncart(c[2]) * ncart(c[3])
    return jk_split(*c) <= 6 and (nf <= 120 or (nf <= 180 and nket <= 9))


def use_row_m(c):
    """Row classes that also get the kernel with only the A rows of K' in LDS
    (the B rows' updates as memory atomics): f-shell bras with a second
    shell above s, whose full rows pass a workgroup's 160 KB from about 1500
    Cartesian functions (13 to 20 rows) while their 10 A rows fit to 2000,
    cut into at most two slices.  Measured on MI355X (def2-TZVP chain, 1442
    AOs, ms per launch, this kernel / the fallback): (fp|ps) 38.8 / 116.5
    cooperative; (fp|ds) 70.6 / 50.7 and (fd|ps) 67.3 / 53.6 (five slices
    each) stay on the fallback."""
    return use_row(c) and c[0] == 3 and c[1] >= 1 and jk_split(*c) <= 2


def use_coop(c):
    """The cooperative LDS kernel (one quartet per block) runs every class
    without a row kernel, and is the fallback of the row classes above 81
    components when their K' rows do not fit in LDS; the tile kernel is the
    fallback of the smaller row classes."""
    return not use_row(c) or nfunc(c) > 81


# Merged row kernels: a generally contracted s shell with two contractions is
# kept whole in merged pairs, (s2 s1) and (s2 s2) for two s shells, (x s2) for
# l(x) = 1, 2; one kernel per quartet class and contraction pattern, up to d
# shells and at most GC_MAX_OUT contracted integrals per lane (a class with a
# larger pattern runs its split pairs on the plain kernels).
GC_LMAX = 2
GC_MAX_OUT = 54
def pair_types(la, lb):
    """Contraction patterns (ca, cb) of the pair groups of a pair class."""
    if lb != 0 or la > GC_LMAX:
        return [(1, 1)]
    return [(1, 1), (2, 1), (2, 2)] if la == 0 else [(1, 1), (1, 2)]


def gc_kernels():
    """[(la, lb, lc, ld, ca, cb, cc, cd)] merged row kernels to instantiate:
    every pattern of a quartet class, for the classes whose patterns all fit
    (a class runs all its merged pair groups or none)."""
    out = []
    for c in quartet_classes(GC_LMAX):
        bra, ket = pair_types(c[0], c[1]), pair_types(c[2], c[3])
        ks = []
        for i, tb in enumerate(bra):
            for j, tk in enumerate(ket):
                if tb == (1, 1) and tk == (1, 1):
                    continue
                if c[:2] == c[2:] and j > i:      # same pair class: ket group <= bra group
                    continue
                ks.append(c + tb + tk)
        if ks and all(k[4] * k[5] * k[6] * k[7] * nfunc(c) <= GC_MAX_OUT for k in ks):
            out.extend(ks)
    return out


def gc_key(k):
    la, lb, lc, ld, ca, cb, cc, cd = k
    return ((((((la * 4 + lb) * 4 + lc) * 4 + ld) * 2 + ca - 1) * 2 + cb - 1) * 2 + cc - 1) * 2 + cd - 1


ROW_TCTL = 8   # ints per class of the persistent row launches (jk_common.h): count + 7 slice counters


def generate():
    os.makedirs(BUILD, exist_ok=True)
    for c in quartet_classes():
        if use_row(c) and jk_split(*c) > ROW_TCTL - 1:
            raise RuntimeError('row class %s has more slices than persistent-launch counters' % (c,))
    files = []
    head = '// generated by pyscf_amd/build.py\n#include "jk_device.h"\nnamespace mjk {\n'
    tail = '}  // namespace mjk\n'
    classes = quartet_classes()
    for c in classes:
        split = jk_split(*c)
        if use_coop(c):
            body = ('int launch_%s(int nb, const JKArgs& a, hipStream_t st)\n{\n'
                    '    return launch_jk_coop<%d, %d, %d, %d>(nb, a, st);\n}\n' % ((_name(c),) + c))
            files.append(_write('%s_coop.hip' % _name(c), head + body + tail))
        else:
            for u0 in range(0, split, SLICES_PER_UNIT):
                body = ''
                for sl in range(u0, min(split, u0 + SLICES_PER_UNIT)):
                    body += ('int launch_%s_%d(int nb, const JKArgs& a, hipStream_t st)\n{\n'
                             '    return launch_jk_slice<%d, %d, %d, %d, %d>(nb, a, st);\n}\n'
                             % ((_name(c), sl) + c + (sl,)))
                files.append(_write('%s_u%d.hip' % (_name(c), u0), head + body + tail))
        if use_row(c):
            # Coulomb rows (_r) and range-separated rows (_q, a.omega != 0)
            for tag, rsh in (('r', 'false'), ('q', 'true')):
                for u0 in range(0, split, SLICES_PER_UNIT):
                    body = ''
                    for sl in range(u0, min(split, u0 + SLICES_PER_UNIT)):
                        body += ('int launch_row%s_%s_%d(int nb, const JKArgs& a, size_t shmem, hipStream_t st)\n'
                                 '{\n    return launch_jk_row<%d, %d, %d, %d, %d, %s>(nb, a, shmem, st);\n}\n'
                                 % ((tag, _name(c), sl) + c + (sl, rsh)))
                    files.append(_write('%s_%s%d.hip' % (_name(c), tag, u0), head + body + tail))
    for c in classes:
        if use_row_m(c):
            split = jk_split(*c)
            for u0 in range(0, split, SLICES_PER_UNIT):
                body = ''
                for sl in range(u0, min(split, u0 + SLICES_PER_UNIT)):
                    body += ('int launch_rowm_%s_%d(int nb, const JKArgs& a, size_t shmem, hipStream_t st)\n'
                             '{\n    return launch_jk_row_m<%d, %d, %d, %d, %d>(nb, a, shmem, st);\n}\n'
                             % ((_name(c), sl) + c + (sl,)))
                files.append(_write('%s_m%d.hip' % (_name(c), u0), head + body + tail))
    gck = gc_kernels()
    bybase = {}
    for k in gck:
        bybase.setdefault(k[:4], []).append(k)
    for c, ks in sorted(bybase.items()):
        body = ''
        for k in ks:
            body += ('int launch_rowg_%d(int nb, const JKArgs& a, size_t shmem, hipStream_t st)\n{\n'
                     '    return launch_jk_row_gc<%d, %d, %d, %d, %d, %d, %d, %d>(nb, a, shmem, st);\n}\n'
                     % ((gc_key(k),) + k))
        files.append(_write('%s_g.hip' % _name(c), head + body + tail))
    decl = ''
    impl = ''
    for k in gck:
        decl += 'int launch_rowg_%d(int nb, const JKArgs& a, size_t shmem, hipStream_t st);\n' % gc_key(k)
    for c in classes:
        split = jk_split(*c)
        if use_coop(c):
            decl += 'int launch_%s(int nb, const JKArgs& a, hipStream_t st);\n' % _name(c)
        else:
            for sl in range(split):
                decl += 'int launch_%s_%d(int nb, const JKArgs& a, hipStream_t st);\n' % (_name(c), sl)
            impl += 'static int launch_%s(int nb, const JKArgs& a, hipStream_t st)\n{\n    int rc = 0;\n' % _name(c)
            for sl in range(split):
                impl += '    rc |= launch_%s_%d(nb, a, st);\n' % (_name(c), sl)
            impl += '    return rc;\n}\n'
        if use_row(c):
            for tag in ('r', 'q'):
                for sl in range(split):
                    decl += ('int launch_row%s_%s_%d(int nb, const JKArgs& a, size_t shmem, hipStream_t st);\n'
                             % (tag, _name(c), sl))
            impl += ('static int launch_row_%s(int nb, const JKArgs& a, size_t shmem, hipStream_t st)\n{\n'
                     '    int rc = 0;\n    if (a.omega != 0.0) {\n' % _name(c))
            for sl in range(split):
                impl += '        rc |= launch_rowq_%s_%d(nb, a, shmem, st);\n' % (_name(c), sl)
            impl += '    } else {\n'
            for sl in range(split):
                impl += '        rc |= launch_rowr_%s_%d(nb, a, shmem, st);\n' % (_name(c), sl)
            impl += '    }\n    return rc;\n}\n'
        if use_row_m(c):
            for sl in range(split):
                decl += ('int launch_rowm_%s_%d(int nb, const JKArgs& a, size_t shmem, hipStream_t st);\n'
                         % (_name(c), sl))
            impl += ('static int launch_rowm_%s(int nb, const JKArgs& a, size_t shmem, hipStream_t st)\n{\n'
                     '    int rc = 0;\n' % _name(c))
            for sl in range(split):
                impl += '    rc |= launch_rowm_%s_%d(nb, a, shmem, st);\n' % (_name(c), sl)
            impl += '    return rc;\n}\n'
    key = lambda c: ((c[0] * 4 + c[1]) * 4 + c[2]) * 4 + c[3]
    rcases = ''.join('    case %d: return launch_row_%s(nblocks, a, shmem, st);\n' % (key(c), _name(c))
                     for c in classes if use_row(c))
    coop = ''.join('    case %d: return 1;\n' % key(c) for c in classes if use_coop(c))
    avail = ''.join('    case %d: return 1;\n' % key(c) for c in classes if use_row(c))
    mcases = ''.join('    case %d: return launch_rowm_%s(nblocks, a, shmem, st);\n' % (key(c), _name(c))
                     for c in classes if use_row_m(c))
    mavail = ''.join('    case %d: return 1;\n' % key(c) for c in classes if use_row_m(c))
    cases = ''.join('    case %d: return launch_%s(nblocks, a, st);\n' % (key(c), _name(c)) for c in classes)
    gcases = ''.join('    case %d: return launch_rowg_%d(nblocks, a, shmem, st);\n' % (gc_key(k), gc_key(k))
                     for k in gck)
    gavail = ''.join('    case %d: return 1;\n' % gc_key(k) for k in gck)
    gsw = ('    switch ((((((((la * 4 + lb) * 4 + lc) * 4 + ld) * 2 + ca - 1) * 2 + cb - 1) * 2 + cc - 1) * 2 + cd - 1)'
           ') {\n')
    disp = ('// generated by pyscf_amd/build.py\n#include "jk_common.h"\nnamespace mjk {\n' + decl + impl +
            'int launch_jk(int la, int lb, int lc, int ld, int nblocks, const JKArgs& a, hipStream_t st)\n{\n'
            '    switch (((la * 4 + lb) * 4 + lc) * 4 + ld) {\n' + cases +
            '    default: return -1;\n    }\n}\n'
            'int coop_kernel(int la, int lb, int lc, int ld)\n{\n'
            '    switch (((la * 4 + lb) * 4 + lc) * 4 + ld) {\n' + coop +
            '    default: return 0;\n    }\n}\n'
            'int row_kernel_available(int la, int lb, int lc, int ld)\n{\n'
            '    switch (((la * 4 + lb) * 4 + lc) * 4 + ld) {\n' + avail +
            '    default: return 0;\n    }\n}\n'
            'int launch_jk_rows(int la, int lb, int lc, int ld, int nblocks, const JKArgs& a, size_t shmem, '
            'hipStream_t st)\n{\n'
            '    switch (((la * 4 + lb) * 4 + lc) * 4 + ld) {\n' + rcases +
            '    default: return -1;\n    }\n}\n'
            'int row_m_available(int la, int lb, int lc, int ld)\n{\n'
            '    switch (((la * 4 + lb) * 4 + lc) * 4 + ld) {\n' + mavail +
            '    default: return 0;\n    }\n}\n'
            'int launch_jk_rows_m(int la, int lb, int lc, int ld, int nblocks, const JKArgs& a, size_t shmem, '
            'hipStream_t st)\n{\n'
            '    switch (((la * 4 + lb) * 4 + lc) * 4 + ld) {\n' + mcases +
            '    default: return -1;\n    }\n}\n'
            'int row_gc_available(int la, int lb, int lc, int ld, int ca, int cb, int cc, int cd)\n{\n'
            '    if (ca < 1 || ca > 2 || cb < 1 || cb > 2 || cc < 1 || cc > 2 || cd < 1 || cd > 2) return 0;\n'
            + gsw + gavail + '    default: return 0;\n    }\n}\n'
            'int launch_jk_rows_gc(int la, int lb, int lc, int ld, int ca, int cb, int cc, int cd, int nblocks, '
            'const JKArgs& a, size_t shmem, hipStream_t st)\n{\n'
            '    if (ca < 1 || ca > 2 || cb < 1 || cb > 2 || cc < 1 || cc > 2 || cd < 1 || cd > 2) return -1;\n'
            + gsw + gcases + '    default: return -1;\n    }\n}\n'
            'int lmax_kernels() { return %d; }\n' % LMAX + tail)
    files.append(_write('jk_dispatch.hip', disp))
    files.extend(generate_df())
    return files


DF_LMAX_AO = 3
DF_LMAX_AUX = 4


def generate_df():
    """Three-centre integral kernels, one unit per auxiliary angular momentum:
    (ij|P) for li >= lj <= 3 and (P|Q) (li <= 4 with the unit function)."""
    files = []
    cases = ''
    for lk in range(DF_LMAX_AUX + 1):
        body = ''
        sw = ''
        for li in range(DF_LMAX_AO + 1):
            for lj in range(li + 1):
                sw += ('    if (!unitj && li == %d && lj == %d) return launch_int3c_t<%d, %d, %d, false>(a, st);\n'
                       % (li, lj, li, lj, lk))
        for li in range(DF_LMAX_AUX + 1):
            sw += ('    if (unitj && li == %d && lj == 0) return launch_int3c_t<%d, 0, %d, true>(a, st);\n'
                   % (li, li, lk))
        body = ('int launch_int3c_l%d(int li, int lj, bool unitj, const Int3cArgs& a, hipStream_t st)\n{\n'
                '%s    return -1;\n}\n' % (lk, sw))
        files.append(_write('df_int3c_l%d.hip' % lk,
                            '// generated by pyscf_amd/build.py\n#include "df_device.h"\nnamespace mjk {\n'
                            + body + '}  // namespace mjk\n'))
        cases += '    case %d: return launch_int3c_l%d(li, lj, unitj, a, st);\n' % (lk, lk)
    decl = ''.join('int launch_int3c_l%d(int li, int lj, bool unitj, const Int3cArgs& a, hipStream_t st);\n' % lk
                   for lk in range(DF_LMAX_AUX + 1))
    disp = ('// generated by pyscf_amd/build.py\n#include "df_device.h"\nnamespace mjk {\n' + decl +
            'int launch_int3c(int li, int lj, int lk, bool unitj, const Int3cArgs& a, hipStream_t st)\n{\n'
            '    switch (lk) {\n' + cases + '    default: return -1;\n    }\n}\n}  // namespace mjk\n')
    files.append(_write('df_dispatch.hip', disp))
    return files


def _write(name, text):
    path = os.path.join(BUILD, name)
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, 'w') as f:
            f.write(text)
    return path


def _deps_mtime():
    return max(os.path.getmtime(os.path.join(CSRC, f)) for f in os.listdir(CSRC)
               if f.endswith(('.h', '.inc', '.hip')))


def _compile(src, force):
    obj = os.path.join(BUILD, os.path.splitext(os.path.basename(src))[0] + '.o')
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= max(_deps_mtime(), os.path.getmtime(src)):
        return obj
    cmd = [HIPCC] + FLAGS + ['-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s' % (src, r.stderr[-4000:]))
    return obj


def build(force=True, verbose=True, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in STATIC_SOURCES] + generate()
    jobs = jobs or min(16, os.cpu_count() or 4)
    # longest (highest angular momentum) units first
    def cost(path):
        b = os.path.basename(path)
        return sum(int(ch) for ch in b[3:7]) if b.startswith('jk_') and b[3:7].isdigit() else 0
    srcs.sort(key=lambda s: -cost(s))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    tmp = OUT + '.tmp'
    cmd = [HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC', '-o', tmp] + objs + LIBS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n%s' % r.stderr[-4000:])
    os.replace(tmp, OUT)
    if verbose:
        print('built', OUT)
    return OUT


if __name__ == '__main__':
    build(force='--force' in sys.argv)
    sys.exit(0)
