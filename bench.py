"""Benchmark of the direct J/K build (get_jk) on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload c2] [--no-cpu]

Workload (default "c2"): (H2O)16 on a jittered 3.0 Angstrom cubic lattice,
def2-SVP (384 AOs), RHF converged to 1e-10 Eh with the GPU engine; then one
step = one J/K build (J and K of the converged density, hermi=1,
direct_scf_tol=1e-13) with the density already resident in HBM.  Every
step is synchronised so that its latency is a sample of the p50/p99.

Printed: one JSON line with value = JK builds/s (whole job, one GPU),
quartets/s, p50/p99 latency, the roofline of the dominant quartet-class
kernel (algorithmic FP64 FLOPs / its HIP-event duration on the launch
stream, against 78.6 TFLOP/s) and the CPU baseline (the oracle's OpenMP
restatement of the reference's direct J/K loop, timed on a bounded sample of
bra shell pairs of the same build and scaled to builds/s).
"""

import argparse
import ctypes
import json
import os
import sys
import time

import numpy

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'JK builds/s + ERI quartets/s, mixed vs FP64, at <=1e-8 Eh; p50/p99 latency'
FP64_PEAK = 78.6e12


def progress(msg):
    """One progress line on stderr (the JSON line alone goes to stdout)."""
    print('[bench %.0fs] %s' % (time.perf_counter() - _T0, msg), file=sys.stderr, flush=True)


_T0 = time.perf_counter()


def kernel_name(cls, kind, rsh=False):
    """Demangled name (as rocprofv3 prints it, less the argument list) of the
    kernel a quartet class ran on: 0 tile, 1 row, 2 cooperative; classes cut
    into slices run one kernel per slice and are named by slice 0.  rsh: the
    row-kernel variant of the range-separated operators."""
    from pyscf_amd.build import jk_split
    la, lb, lc, ld = cls
    if kind == 2:
        return 'mjk::k_jk_coop<%d, %d, %d, %d>' % cls
    if kind == 1:
        return 'mjk::k_jk_row<%d, %d, %d, %d, %d, 0, %s>' % (la, lb, lc, ld, jk_split(la, lb, lc, ld),
                                                             'true' if rsh else 'false')
    return 'mjk::k_jk<%d, %d, %d, %d, %d, 0>' % (la, lb, lc, ld, jk_split(la, lb, lc, ld))


def traffic_bytes(kernel):
    """HBM bytes per launch of `kernel` on the c2 workload, from the calibrated
    FETCH_SIZE and WRITE_SIZE counter runs summarised in profiles/ (None if
    the kernel has no entry there)."""
    path = os.path.join(ROOT, 'profiles', 'jk_c2_traffic.json')
    try:
        with open(path) as f:
            rec = json.load(f)['kernels'][kernel]
    except (OSError, KeyError, ValueError):
        return None
    return rec['fetch_bytes_per_launch'] + rec['write_bytes_per_launch']


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--workload', default='c2', choices=['c1', 'c2', 'c3', 'c4', 'c5'])
    ap.add_argument('--nmol', type=int, default=256, help='c4: molecules in the batch')
    ap.add_argument('--nreq', type=int, default=48, help='c5: requests in the stream')
    ap.add_argument('--rate', type=float, default=20.0, help='c5: Poisson arrival rate (1/s)')
    ap.add_argument('--max-ao', type=int, default=1500, help='c5: largest request (AOs)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--dm-out', default=None, help='c2: save the converged density (.npy)')
    ap.add_argument('--dm-in', default=None,
                    help='c2: load the density from a .npy written by --dm-out instead of running the SCF '
                         '(profiling runs: every J/K launch is then one of the converged density)')
    ap.add_argument('--cpu-seconds', type=float, default=10.0)
    ap.add_argument('--fp32-tol', type=float, default=1e-7)
    ap.add_argument('--omega', type=float, default=0.0,
                    help='c1-c3: time J/K builds of the range-separated operator (> 0 erf, < 0 erfc) '
                         'on the converged Coulomb density')
    return ap.parse_args()


def workload(name):
    from pyscf_amd import gto, workloads
    if name == 'c1':
        return gto.M(atom=workloads.h2o(), basis='cc-pvdz'), 'H2O cc-pVDZ (24 AOs)'
    if name == 'c2':
        return (gto.M(atom=workloads.water_cluster(16, seed=1), basis='def2-svp'),
                '(H2O)16 jittered 3.0 A lattice (seed 1), def2-SVP (384 AOs)')
    return gto.M(atom=workloads.alkane(30), basis='cc-pvdz'), 'C30H62 all-trans, cc-pVDZ (730 AOs)'


def cpu_baseline(mol, dm, tol, seconds):
    """Oracle (OpenMP C) on a bounded sample of bra shell pairs."""
    from oracle import oracle as O
    ncores = len(os.sched_getaffinity(0))
    O.set_threads(ncores)
    opt = O.VHFOpt(mol, tol)
    dms = dm[None]
    _, _, n_total = O.direct_full(mol, dms, opt, pair_stride=0)
    t = time.perf_counter()
    _, _, n_probe = O.direct_full(mol, dms, opt, pair_stride=97)
    rate = n_probe / max(time.perf_counter() - t, 1e-6)
    stride = max(1, int(n_total / max(rate * seconds, 1)))
    t = time.perf_counter()
    _, _, n_s = O.direct_full(mol, dms, opt, pair_stride=stride)
    dt = time.perf_counter() - t
    qps = n_s / dt
    return {'value': qps / n_total, 'unit': 'JK builds/s', 'cores': ncores, 'kind': 'port',
            'quartets_per_s': qps,
            'sample': 'bra shell pairs ij with ij %% %d == 0 of the same screened build: %d of %d '
                      'unique shell quartets in %.2f s, scaled by quartet count' % (stride, n_s, n_total, dt)}


def bench_batched(args):
    """256 random-walk chains (def2-SVP): batched SCF (one launch per quartet
    class and iteration for all molecules) vs the same J/K builds issued one
    molecule at a time."""
    from pyscf_amd import gto, workloads
    from pyscf_amd.scf import _vhf, batch
    mols = [gto.M(atom=workloads.random_chain([4, i]), basis='def2-svp') for i in range(args.nmol)]
    naos = [m.nao_nr() for m in mols]
    res = {}
    for label, tol32 in (('fp64', 0.0), ('mixed', args.fp32_tol)):
        progress('c4 %s: batched RHF of %d molecules' % (label, len(mols)))
        b = batch.BatchRHF(mols, fp32_tol=tol32, log=lambda s, label=label: progress('c4 %s %s' % (label, s)))
        b.conv_tol = 1e-9
        t0 = time.perf_counter()
        e = b.kernel()
        t_scf = time.perf_counter() - t0
        t_jk = sum(b.build_times)
        res[label] = {'e_tot': e, 'scf_s': t_scf, 'iterations': b.cycles, 'jk_s': t_jk,
                      'converged': int(b.converged.sum()), 'conv_mask': numpy.array(b.converged),
                      'mol_builds_per_s': len(mols) * len(b.build_times) / t_jk}
    # unbatched baseline: the FP64 J/K build of every molecule, one call each
    dms = b.dms
    eng = _vhf.get_engine()
    progress('c4: unbatched baseline, %d uploads' % len(mols))
    opts = [_vhf.VHFOpt(m, direct_scf_tol=1e-13, engine=eng) for m in mols]
    for o, d in zip(opts, dms):
        o.basis.get_jk(d[None], 1, True, True, 1e-13)
    t0 = time.perf_counter()
    for o, d in zip(opts, dms):
        o.basis.get_jk(d[None], 1, True, True, 1e-13)
    t_single = time.perf_counter() - t0
    single = len(mols) / t_single
    de = numpy.abs(res['mixed']['e_tot'] - res['fp64']['e_tot'])
    both = res['mixed']['conv_mask'] & res['fp64']['conv_mask']
    # molecules whose batched mixed and FP64 energies differ by more than 1e-8:
    # rerun them alone (FP64 and mixed, tighter convergence) to tell a precision
    # error from two SCF runs that settled in different solutions
    from pyscf_amd.scf import rhf
    outliers = []
    for i in [j for j in numpy.argsort(-de) if both[j] and de[j] > 1e-8][:5]:
        progress('c4: rechecking molecule %d alone (|dE| %.2e)' % (i, de[i]))
        es = []
        for t32 in (0.0, args.fp32_tol):
            mf = rhf.RHF(mols[i], direct_scf_tol=1e-13, fp32_tol=t32)
            mf.conv_tol = 1e-10
            es.append(mf.kernel())
        outliers.append({'molecule': int(i), 'batched_dE': float(de[i]), 'alone_fp64': es[0], 'alone_mixed': es[1],
                         'alone_dE': es[1] - es[0], 'batched_fp64': float(res['fp64']['e_tot'][i]),
                         'batched_mixed': float(res['mixed']['e_tot'][i])})
    out = {
        'metric': METRIC, 'value': res['mixed']['mol_builds_per_s'], 'unit': 'molecule JK builds/s',
        'n_gpus': 1, 'steps': res['mixed']['iterations'], 'warmup': 0,
        'ms_per_step': 1e3 * res['mixed']['jk_s'] / res['mixed']['iterations'], 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64+f32', 'data': 'synthetic random-walk chains',
        'config': {'workload': '%d random-walk chains (3-12 heavy atoms, seeds [4,i]), def2-SVP, '
                   '%d-%d AOs, batched RHF to 1e-9' % (len(mols), min(naos), max(naos))},
        'fp64_batched_mol_builds_per_s': res['fp64']['mol_builds_per_s'],
        'mixed_batched_mol_builds_per_s': res['mixed']['mol_builds_per_s'],
        'unbatched_fp64_mol_builds_per_s': single,
        'speedup_mixed_batched_vs_unbatched_fp64': res['mixed']['mol_builds_per_s'] / single,
        'speedup_fp64_batched_vs_unbatched_fp64': res['fp64']['mol_builds_per_s'] / single,
        'max_abs_dE_mixed_vs_fp64': float(de[both].max()) if both.any() else None,
        'max_abs_dE_mixed_vs_fp64_incl_unconverged': float(de.max()),
        'molecules_dE_over_1e-8': int((de[both] > 1e-8).sum()),
        'dE_outliers_rerun_alone': outliers,
        'converged': [res['fp64']['converged'], res['mixed']['converged']],
        'scf_seconds': [res['fp64']['scf_s'], res['mixed']['scf_s']],
    }
    print(json.dumps(out))


def bench_serving(args):
    """Poisson stream of get_jk and RHF requests (50-max AOs, lognormal;
    70% def2-SVP, 30% def2-TZVP) through the continuous-batching server."""
    from pyscf_amd import gto, workloads
    from pyscf_amd.server import JKServer, Request
    rng = numpy.random.default_rng(5)
    reqs = []
    for i in range(args.nreq):
        basis = 'def2-svp' if rng.random() < 0.7 else 'def2-tzvp'
        target = float(numpy.clip(rng.lognormal(numpy.log(200), 0.8), 50, args.max_ao))
        per_heavy = 14 if basis == 'def2-svp' else 31
        nheavy = max(2, int(round(target / (per_heavy + 8))))
        mol = gto.M(atom=workloads.random_chain([5, i], nmin=nheavy, nmax=nheavy), basis=basis)
        kind = 'rhf' if rng.random() < 0.5 else 'get_jk'
        dm = None
        if kind == 'get_jk':
            c = rng.standard_normal((mol.nao_nr(), max(2, mol.nelectron // 2))) / 4
            dm = c @ c.T
        reqs.append((kind, mol, dm, basis))
    gaps = rng.exponential(1.0 / args.rate, size=len(reqs))
    progress('c5: %d requests built' % len(reqs))
    srv = JKServer(nslots=8)
    t0 = time.perf_counter()
    futs = []
    for (kind, mol, dm, _), g in zip(reqs, gaps):
        time.sleep(g)
        futs.append(srv.submit(Request(kind, mol, dm)))
    ok = 0
    for i, f in enumerate(futs):
        try:
            f.result(timeout=1200)
            ok += 1
        except Exception as e:
            progress('c5: request %d failed: %s' % (i, e))
        progress('c5: %d of %d requests done' % (i + 1, len(futs)))
    t_all = time.perf_counter() - t0
    srv.close()
    lat = numpy.array([r.t_done - r.t_arrival for r in srv.finished])
    naos = [m.nao_nr() for _, m, _, _ in reqs]
    out = {
        'metric': METRIC, 'value': ok / t_all, 'unit': 'requests/s', 'n_gpus': 1, 'steps': len(reqs),
        'warmup': 0, 'ms_per_step': 1e3 * t_all / len(reqs), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f64', 'data': 'synthetic random-walk chains',
        'config': {'workload': 'Poisson stream (rate %.1f/s, seed 5) of %d requests, 50/50 get_jk/RHF, '
                   '70%% def2-SVP / 30%% def2-TZVP, %d-%d AOs, 8 slots, 80%% HBM budget'
                   % (args.rate, len(reqs), min(naos), max(naos))},
        'p50_ms': float(numpy.percentile(lat, 50) * 1e3), 'p99_ms': float(numpy.percentile(lat, 99) * 1e3),
        'completed': ok,
    }
    print(json.dumps(out))


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py runs on one GPU only: the J/K engine has no multi-GPU path; '
              'got --gpus %d' % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload == 'c4':
        return bench_batched(args)
    if args.workload == 'c5':
        return bench_serving(args)
    from pyscf_amd import lib
    from pyscf_amd.scf import _vhf, rhf

    mol, desc = workload(args.workload)
    nao = mol.nao_nr()
    tol = 1e-13
    t0 = time.perf_counter()
    if args.dm_in:
        dm = numpy.ascontiguousarray(numpy.load(args.dm_in, allow_pickle=False))
        opt = _vhf.VHFOpt(mol, direct_scf_tol=tol)
        e64 = None
        scf_info = {'skipped': 'density from %s' % os.path.basename(args.dm_in)}
    else:
        mf = rhf.RHF(mol, direct_scf_tol=tol)
        mf.conv_tol = 1e-10
        e64 = mf.kernel()
        dm = numpy.ascontiguousarray(mf.dm)
        opt = mf._opt[None]
        scf_info = {'e_tot': e64, 'cycles': mf.cycles, 'converged': mf.converged,
                    'seconds': time.perf_counter() - t0}
        if args.dm_out:
            numpy.save(args.dm_out, dm)

    if args.omega:
        # bounds and builds of the range-separated operator, on the same density
        with mol.with_range_coulomb(args.omega):
            opt = _vhf.VHFOpt(mol, direct_scf_tol=tol)
    eng = opt.engine
    L = eng.lib
    basis = opt.basis
    ctx = eng.handle
    nbytes = dm.nbytes
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    d_dm, d_vj, d_vk = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    for p in (d_dm, d_vj, d_vk):
        eng.check(L.mjk_device_alloc(ctx, nbytes, ctypes.byref(p)), 'device_alloc')
    eng.check(L.mjk_memcpy_h2d(ctx, d_dm, ptr(dm), nbytes), 'h2d')

    def timed(fp32_tol):
        """warmup, then args.steps synchronised builds with per-class HIP events."""
        st = lib.mjk_stats()
        def build(stats=None):
            rc = L.mjk_get_jk_device(ctx, basis.handle, d_dm, 1, 1, 1, 1, tol, fp32_tol, d_vj, d_vk,
                                     ctypes.byref(stats) if stats is not None else None)
            eng.check(rc, 'mjk_get_jk_device')
        build(st)                      # first call: tile lists for this cutoff, work buffers
        for _ in range(args.warmup):
            build()
        eng.check(L.mjk_synchronize(ctx), 'sync')
        maxc = 64
        cls = numpy.zeros(4 * maxc, dtype=numpy.int32)
        ms = numpy.zeros(maxc)
        nq = numpy.zeros(maxc, dtype=numpy.int64)
        npq = numpy.zeros(maxc, dtype=numpy.int64)
        fl = numpy.zeros(maxc)
        nl = numpy.zeros(maxc, dtype=numpy.int64)
        # work per class from one counted build (untimed) ...
        eng.check(L.mjk_profile(ctx, basis.handle, 1), 'profile')
        build()
        k = L.mjk_profile_read(ctx, basis.handle, maxc, ptr(cls), ptr(ms), ptr(nq), ptr(npq), ptr(fl), ptr(nl))
        work = {tuple(cls[4 * i:4 * i + 4]): (int(nq[i] // nl[i]), fl[i] / nl[i]) for i in range(k)}
        # ... and kernel times from HIP events over the timed builds (no tallies kept)
        eng.check(L.mjk_profile(ctx, basis.handle, 2), 'profile')
        lat = []
        eng.check(L.mjk_synchronize(ctx), 'sync')
        t_start = time.perf_counter()
        for _ in range(args.steps):
            t = time.perf_counter()
            build()
            eng.check(L.mjk_synchronize(ctx), 'sync')
            lat.append(time.perf_counter() - t)
        eng.check(L.mjk_synchronize(ctx), 'sync')
        t_total = time.perf_counter() - t_start
        k = L.mjk_profile_read(ctx, basis.handle, maxc, ptr(cls), ptr(ms), ptr(nq), ptr(npq), ptr(fl), ptr(nl))
        kinds = numpy.zeros(maxc, dtype=numpy.int32)
        L.mjk_profile_kinds(ctx, basis.handle, maxc, ptr(kinds))
        span, nspan = ctypes.c_double(), ctypes.c_longlong()
        eng.check(L.mjk_profile_span(ctx, basis.handle, ctypes.byref(span), ctypes.byref(nspan)), 'profile_span')
        L.mjk_profile(ctx, basis.handle, 0)
        classes = []
        for i in range(k):
            key = tuple(int(x) for x in cls[4 * i:4 * i + 4])
            nq1, fl1 = work.get(key, (0, 0.0))
            t1 = ms[i] / nl[i]
            kname = kernel_name(key, int(kinds[i]), bool(args.omega))
            classes.append({'class': '(%d%d|%d%d)' % key, 'kernel': kname,
                            'ms_per_launch': t1, 'quartets': nq1, 'flops': fl1,
                            'tflops': fl1 / (t1 * 1e-3) / 1e12 if t1 > 0 else 0.0})
        classes.sort(key=lambda c: -c['ms_per_launch'])
        vj = numpy.empty_like(dm)
        vk = numpy.empty_like(dm)
        eng.check(L.mjk_memcpy_d2h(ctx, ptr(vj), d_vj, nbytes), 'd2h')
        eng.check(L.mjk_memcpy_d2h(ctx, ptr(vk), d_vk, nbytes), 'd2h')
        return {'t_total': t_total, 'lat': lat, 'classes': classes, 'stats': st, 'vj': vj, 'vk': vk,
                'span_ms': span.value / max(nspan.value, 1)}

    r64 = timed(0.0)
    rmx = timed(args.fp32_tol) if args.fp32_tol > 0 else None
    for p in (d_dm, d_vj, d_vk):
        L.mjk_device_free(ctx, p)
    # mixed-precision SCF energy
    emx = None
    if rmx is not None and e64 is not None and not args.omega:
        mfm = rhf.RHF(mol, direct_scf_tol=tol, fp32_tol=args.fp32_tol)
        mfm.conv_tol = 1e-10
        emx = mfm.kernel()
    t_total, lat, classes, st = r64['t_total'], r64['lat'], r64['classes'], r64['stats']
    # the classes run concurrently on several streams: the dominant kernel is the
    # one with the most algorithmic FLOPs; its duration is measured while it
    # shares the GPU with the others
    dom = max(classes, key=lambda c: c['flops'])
    span_ms = r64['span_ms']
    ms_step = t_total / args.steps * 1e3
    builds = args.steps / t_total
    kernel_ms = sum(c['ms_per_launch'] for c in classes)
    total_flops = sum(c['flops'] for c in classes)
    out = {
        'metric': METRIC,
        'value': builds,
        'unit': 'JK builds/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms_step,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic geometry; converged RHF density from this engine',
        'config': {'workload': desc + ', one J/K build of the converged RHF density, hermi=1, '
                   'direct_scf_tol=1e-13', 'nao': nao, 'n_dm': 1, 'precision': 'fp64', 'omega': args.omega},
        'quartets_per_s': st.quartets * builds,
        'quartets_per_build': int(st.quartets),
        'p50_ms': float(numpy.percentile(lat, 50) * 1e3),
        'p99_ms': float(numpy.percentile(lat, 99) * 1e3),
        'scf': scf_info,
        'gpu_quartet_phase_ms_per_build': span_ms,
        'sum_of_class_kernel_ms_per_build': kernel_ms,
        'fp64_tflops_all_classes': total_flops / (span_ms * 1e-3) / 1e12,
        # compute-bound against the FP64 peak (78.6 TFLOP/s is both the matrix and the
        # vector FP64 peak of MI355X; the Rys kernels run on the vector ALU)
        'roofline': {'bound': 'mfma', 'compute': 'fp64 vector ALU', 'kernel': dom['kernel'], 'class': dom['class'],
                     'achieved': dom['tflops'], 'peak': FP64_PEAK / 1e12, 'unit': 'TFLOP/s',
                     'frac': dom['tflops'] * 1e12 / FP64_PEAK,
                     'traffic': traffic_bytes(dom['kernel']) if args.workload == 'c2' else None,
                     'flops_per_launch': dom['flops'], 'ms_per_launch': dom['ms_per_launch'],
                     'build': {'what': 'all quartet-class kernels of one J/K build, concurrent on 4 streams',
                               'achieved': total_flops / (span_ms * 1e-3) / 1e12,
                               'frac': total_flops / (span_ms * 1e-3) / FP64_PEAK,
                               'flops': total_flops, 'ms': span_ms}},
        'classes': classes[:8],
    }
    if rmx is not None:
        b_mx = args.steps / rmx['t_total']
        out['mixed'] = {
            'fp32_tol': args.fp32_tol,
            'jk_builds_per_s': b_mx,
            'speedup_vs_fp64': b_mx / builds,
            'quartets_per_s': rmx['stats'].quartets * b_mx,
            'fp32_quartet_fraction': rmx['stats'].quartets_fp32 / max(rmx['stats'].quartets, 1),
            'p50_ms': float(numpy.percentile(rmx['lat'], 50) * 1e3),
            'p99_ms': float(numpy.percentile(rmx['lat'], 99) * 1e3),
            'max_abs_dJ': float(abs(rmx['vj'] - r64['vj']).max()),
            'max_abs_dK': float(abs(rmx['vk'] - r64['vk']).max()),
            'e_tot': emx,
            'dE_vs_fp64': emx - e64 if emx is not None else None,
        }
    if not args.no_cpu:
        with mol.with_range_coulomb(args.omega or None):
            out['cpu_baseline'] = cpu_baseline(mol, dm, tol, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == '__main__':
    main()
