"""Benchmark of the direct J/K build (get_jk) on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload c3] [--no-cpu] [--no-oracle]

Default workload "c3": all-trans C30H62, cc-pVDZ (730 AOs).  RHF is
converged to 1e-10 Eh with the GPU engine; then one step = one J/K build of
the converged density (J and K, hermi=1, direct_scf_tol=1e-13) with the
density already resident in HBM.  Every step is synchronised, so that its
latency is a sample of the p50/p99.

The JSON line (stdout) carries:
  * value = FP64 JK builds/s (one GPU), quartets/s, p50/p99;
  * mixed: the same builds with FP32 recurrences for quartets bounded below
    fp32_tol = 1e-7 (builds/s, speed-up) and the mixed-precision RHF energy
    against the FP64 one;
  * accuracy: max |dJ|, max |dK| of the FP64 and the mixed builds against
    one oracle build of the same density (the oracle's OpenMP restatement of
    the reference's direct J/K loop, on all host cores);
  * roofline: the quartet-class kernel with the most algorithmic FP64 FLOPs,
    its FLOPs per launch over its average duration from HIP events on its
    launch stream during the timed builds, against 78.6 TFLOP/s (the kernels
    run on the FP64 vector ALU); also its average over every launch of the
    process (what a kernel-trace average of the same run reports);
  * cpu_baseline: the oracle timed on a bounded sample of bra shell pairs of
    the same screened build (1 warm-up, median of 5), scaled by quartet count;
  * secondary: the (H2O)16 def2-SVP build (384 AOs) with its oracle check.

Other modes: --workload c1/c2 (single molecules), c4 (256 chains batched,
active-molecule builds/s against the same molecules' own RHF runs), c5
(serving stream), c3df (density-fitted J/K of C30H62 with cc-pVDZ-JKFIT); --pmc-run DM.npy (two builds of a saved density and a
64 MiB calibration copy, for rocprofv3 --pmc runs); --pmc-summary FETCH_DIR
WRITE_DIR (per-kernel HBM bytes per launch from those runs' CSVs).
"""

import argparse
import ctypes
import glob
import json
import os
import subprocess
import sys
import time

import numpy

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'JK builds/s + ERI quartets/s, mixed vs FP64, at <=1e-8 Eh; p50/p99 latency'
FP64_PEAK = 78.6e12
TOL = 1e-13
_T0 = time.perf_counter()
# host cores of this process, read before OpenMP binds the main thread to one of them
NCORES = len(os.sched_getaffinity(0))


def progress(msg):
    """One progress line on stderr (the JSON line alone goes to stdout)."""
    print('[bench %.0fs] %s' % (time.perf_counter() - _T0, msg), file=sys.stderr, flush=True)


def kernel_name(cls, kind, rsh=False, ctr=(1, 1, 1, 1)):
    """Demangled name (as rocprofv3 prints it, less the argument list) of the
    kernel a quartet class ran on: 0 tile, 1 row, 2 cooperative, 3 merged row
    kernel (contraction counts ctr), 4 row kernel with the B rows of K' in
    memory; classes cut into slices run one kernel per slice and are named by
    slice 0.  rsh: the row-kernel variant of the range-separated operators."""
    from pyscf_amd.build import jk_split
    la, lb, lc, ld = cls
    if kind == 4:
        return 'mjk::k_jk_row_m<%d, %d, %d, %d, %d, 0>' % (la, lb, lc, ld, jk_split(la, lb, lc, ld))
    if kind == 3:
        return 'mjk::k_jk_row_gc<%d, %d, %d, %d, %d, %d, %d, %d>' % (tuple(cls) + tuple(ctr))
    if kind == 2:
        return 'mjk::k_jk_coop<%d, %d, %d, %d>' % cls
    if kind == 1:
        return 'mjk::k_jk_row<%d, %d, %d, %d, %d, 0, %s>' % (la, lb, lc, ld, jk_split(la, lb, lc, ld),
                                                             'true' if rsh else 'false')
    return 'mjk::k_jk<%d, %d, %d, %d, %d, 0>' % (la, lb, lc, ld, jk_split(la, lb, lc, ld))


def traffic_bytes(kernel, workload):
    """HBM bytes per launch of `kernel` from the calibrated FETCH_SIZE and
    WRITE_SIZE summary of this workload in profiles/ (None without one)."""
    path = os.path.join(ROOT, 'profiles', 'jk_%s_traffic.json' % workload)
    try:
        with open(path) as f:
            kernels = json.load(f)['kernels']
        rec = kernels[kernel] if kernel in kernels else kernels['void ' + kernel]   # the trace prints 'void '
    except (OSError, KeyError, ValueError):
        return None
    return rec['fetch_bytes_per_launch'] + rec['write_bytes_per_launch']


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--workload', default='c3', choices=['c1', 'c2', 'c3', 'c4', 'c5', 'c3df'])
    ap.add_argument('--nmol', type=int, default=256, help='c4: molecules in the batch')
    ap.add_argument('--nreq', type=int, default=48, help='c5: requests in the stream')
    ap.add_argument('--rate', type=float, default=20.0, help='c5: Poisson arrival rate (1/s)')
    ap.add_argument('--max-ao', type=int, default=1500, help='c5: largest request (AOs)')
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline')
    ap.add_argument('--no-oracle', action='store_true', help='skip the oracle accuracy check')
    ap.add_argument('--no-secondary', action='store_true', help='skip the (H2O)16 secondary line')
    ap.add_argument('--dm-out', default=None, help='save the converged density (.npy)')
    ap.add_argument('--dm-in', default=None,
                    help='load the density from a .npy written by --dm-out instead of running the SCF '
                         '(profiling runs: every J/K launch is then one of the converged density)')
    ap.add_argument('--cpu-seconds', type=float, default=3.0, help='CPU baseline: seconds per timed sample')
    ap.add_argument('--fp32-tol', type=float, default=1e-7)
    ap.add_argument('--omega', type=float, default=0.0,
                    help='c1-c3: time J/K builds of the range-separated operator (> 0 erf, < 0 erfc) '
                         'on the converged Coulomb density')
    ap.add_argument('--pmc-run', default=None, metavar='DM.npy',
                    help='counter runs: two builds of the saved density, then a 64 MiB calibration copy')
    ap.add_argument('--pmc-summary', nargs=2, default=None, metavar=('FETCH_DIR', 'WRITE_DIR'),
                    help='HBM bytes per launch and kernel from the FETCH_SIZE and WRITE_SIZE runs')
    ap.add_argument('--class-reps', type=int, default=20,
                    help='c1-c3: timed launches of every quartet class alone (0: skip the per-class times)')
    ap.add_argument('--cpu-child', default=None, metavar='NPZ',
                    help='internal: the CPU side (oracle, CPU baseline) of a single-molecule run')
    return ap.parse_args()


def workload(name):
    from pyscf_amd import gto, workloads
    if name == 'c1':
        return gto.M(atom=workloads.h2o(), basis='cc-pvdz'), 'H2O cc-pVDZ (24 AOs)'
    if name == 'c2':
        return (gto.M(atom=workloads.water_cluster(16, seed=1), basis='def2-svp'),
                '(H2O)16 jittered 3.0 A lattice (seed 1), def2-SVP (384 AOs)')
    return gto.M(atom=workloads.alkane(30), basis='cc-pvdz'), 'C30H62 all-trans, cc-pVDZ (730 AOs)'


def cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def oracle_module():
    """The oracle, on every host core with threads bound close (set before its
    OpenMP runtime starts), with interpolated roots."""
    os.environ.setdefault('OMP_PROC_BIND', 'close')
    from oracle import oracle as O
    O.set_threads(NCORES)
    O.set_roots(1)
    return O


def oracle_jk(mol, dm):
    O = oracle_module()
    t = time.perf_counter()
    opt = O.VHFOpt(mol, TOL)
    j0, k0 = O.get_jk(mol, dm, hermi=1, vhfopt=opt)
    return j0, k0, time.perf_counter() - t


def cpu_baseline(mol, dm, seconds):
    """The oracle on a bounded sample of bra shell pairs of the same screened
    build: stride chosen for about `seconds` per run, 1 warm-up, median of 5,
    scaled to whole builds by the quartet count."""
    O = oracle_module()
    ncores = NCORES
    opt = O.VHFOpt(mol, TOL)
    dms = dm[None]
    _, _, n_total = O.direct_full(mol, dms, opt, pair_stride=0)
    t = time.perf_counter()
    _, _, n_probe = O.direct_full(mol, dms, opt, pair_stride=97)
    rate = n_probe / max(time.perf_counter() - t, 1e-6)
    stride = max(1, int(n_total / max(rate * seconds, 1)))
    O.direct_full(mol, dms, opt, pair_stride=stride)           # warm-up
    runs = []
    for _ in range(5):
        t = time.perf_counter()
        _, _, n_s = O.direct_full(mol, dms, opt, pair_stride=stride)
        runs.append(time.perf_counter() - t)
    dt = float(numpy.median(runs))
    qps = n_s / dt
    return {'value': qps / n_total, 'unit': 'JK builds/s', 'cores': ncores, 'kind': 'port',
            'cpu_model': cpu_model(), 'quartets_per_s': qps, 'runs_s': runs,
            'sample': 'bra shell pairs ij with ij %% %d == 0 of the same screened build (%d of %d unique shell '
                      'quartets), oracle with interpolated Rys roots, OMP_PROC_BIND=close, 1 warm-up, median of '
                      '5 runs %.2f s, scaled by quartet count' % (stride, n_s, n_total, dt)}


def cpu_child(args):
    """CPU side of a single-molecule benchmark, in a process of its own that
    never opens the GPU: the oracle build of the density the GPU used (max
    |dJ|, |dK| of the GPU results against it) and the CPU baseline.  Reads the
    .npz the GPU process wrote; prints one JSON line."""
    data = numpy.load(args.cpu_child, allow_pickle=False)
    name = str(data['workload'])
    omega = float(data['omega'])
    mol, _ = workload(name)
    dm = numpy.ascontiguousarray(data['dm'])
    out = {}
    if not args.no_oracle:
        progress('%s: oracle build of the same density (CPU process)' % name)
        with mol.with_range_coulomb(omega or None):
            j0, k0, t_orc = oracle_jk(mol, dm)
        acc = {'reference': 'oracle (CPU restatement of the reference direct J/K), same density and screening',
               'oracle_build_s': t_orc,
               'fp64_max_abs_dJ': float(abs(data['vj64'] - j0).max()),
               'fp64_max_abs_dK': float(abs(data['vk64'] - k0).max())}
        if 'vjmx' in data:
            acc['mixed_max_abs_dJ'] = float(abs(data['vjmx'] - j0).max())
            acc['mixed_max_abs_dK'] = float(abs(data['vkmx'] - k0).max())
        out['accuracy'] = acc
    if not args.no_cpu:
        progress('%s: CPU baseline (CPU process)' % name)
        with mol.with_range_coulomb(omega or None):
            out['cpu_baseline'] = cpu_baseline(mol, dm, args.cpu_seconds)
    print(json.dumps(out))


def run_cpu_child(args, name, omega, dm, r64, rmx, with_oracle, with_cpu):
    """Writes the density and the GPU J/K, runs the CPU child, returns its JSON."""
    if not (with_oracle or with_cpu):
        return {}
    os.makedirs(os.path.join(ROOT, 'gpurun_out'), exist_ok=True)
    path = os.path.join(ROOT, 'gpurun_out', 'bench_%s_%d.npz' % (name, os.getpid()))
    arrays = {'workload': numpy.array(name), 'omega': numpy.array(omega), 'dm': dm,
              'vj64': r64['vj'], 'vk64': r64['vk']}
    if rmx is not None:
        arrays['vjmx'], arrays['vkmx'] = rmx['vj'], rmx['vk']
    numpy.savez(path, **arrays)
    cmd = [sys.executable, os.path.abspath(__file__), '--cpu-child', path, '--cpu-seconds', str(args.cpu_seconds)]
    if not with_oracle:
        cmd.append('--no-oracle')
    if not with_cpu:
        cmd.append('--no-cpu')
    res = subprocess.run(cmd, stdout=subprocess.PIPE, check=True)
    os.remove(path)
    return json.loads(res.stdout.decode().strip().splitlines()[-1])


def timed_builds(basis, fp32_tol, steps, warmup):
    """Warm-up builds, then `steps` plain builds of the resident density (each
    a graph replay, synchronised, J and K left in HBM): wall time per build."""
    for _ in range(warmup):
        basis.get_jk_resident(1, 1, True, True, TOL, fp32_tol)
    lat = []
    t_start = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        basis.get_jk_resident(1, 1, True, True, TOL, fp32_tol)
        lat.append(time.perf_counter() - t)
    return time.perf_counter() - t_start, lat


def class_times(basis, fp32_tol, reps, omega):
    """Each quartet class alone: builds restricted to it (class filter) with
    statistics, whose HIP events on the build stream bracket exactly that
    class's kernel launch (the class runs on the build stream when it is the
    only one); one untimed build first, then the mean over `reps`."""
    out = []
    for k, l4, ctr, nblocks, kern in basis.classes():
        if nblocks == 0 or kern < 0:     # no task, or the other representation of its quartets runs
            continue
        basis.set_class_filter(k)
        st = basis.get_jk_resident(1, 1, True, True, TOL, fp32_tol, stats=True)
        ms = []
        for _ in range(reps):
            ms.append(basis.get_jk_resident(1, 1, True, True, TOL, fp32_tol, stats=True)['ms_quartets'])
        t1 = float(numpy.mean(ms))
        label = '(%d%d|%d%d)' % l4 if ctr == (1, 1, 1, 1) else '(%d%d|%d%d) ctr %d%d%d%d' % (l4 + ctr)
        out.append({'class': label, 'kernel': kernel_name(l4, kern, bool(omega), ctr),
                    'ms_per_launch': t1, 'launches': reps, 'quartets': int(st['quartets']),
                    'quartets_fp32': int(st['quartets_fp32']), 'flops': st['flops'],
                    'tflops': st['flops'] / (t1 * 1e-3) / 1e12 if t1 > 0 else 0.0})
    basis.set_class_filter(-1)
    out.sort(key=lambda c: -c['flops'])
    return out


def gpu_leg(basis, dm, fp32_tol, steps, warmup, class_reps, omega):
    """Plain timed builds, one counted build (work and GPU phase times) and the
    per-class timings of one precision; J and K of the density."""
    vj, vk = basis.get_jk(dm[None], 1, True, True, TOL, fp32_tol)      # uploads the density
    t_total, lat = timed_builds(basis, fp32_tol, steps, warmup)
    st = basis.get_jk_resident(1, 1, True, True, TOL, fp32_tol, stats=True)
    classes = class_times(basis, fp32_tol, class_reps, omega) if class_reps > 0 else []
    return {'t_total': t_total, 'lat': lat, 'stats': st, 'classes': classes,
            'vj': vj.reshape(dm.shape), 'vk': vk.reshape(dm.shape)}


def single_molecule(args, name, steps, warmup, with_cpu, with_oracle, fp32_tol, class_reps, primary=True):
    """SCF, timed FP64 (and mixed) builds, per-class kernel times, then the
    oracle accuracy and the CPU baseline in a CPU-only child process."""
    from pyscf_amd.scf import _vhf, rhf
    mol, desc = workload(name)
    nao = mol.nao_nr()
    t0 = time.perf_counter()
    if args.dm_in and primary:
        dm = numpy.ascontiguousarray(numpy.load(args.dm_in, allow_pickle=False))
        opt = _vhf.VHFOpt(mol, direct_scf_tol=TOL)
        e64 = None
        scf_info = {'skipped': 'density from %s' % os.path.basename(args.dm_in)}
    else:
        progress('%s: FP64 RHF' % name)
        mf = rhf.RHF(mol, direct_scf_tol=TOL)
        mf.conv_tol = 1e-10
        e64 = mf.kernel()
        dm = numpy.ascontiguousarray(mf.dm)
        opt = mf._opt[None]
        scf_info = {'e_tot': e64, 'cycles': mf.cycles, 'converged': mf.converged,
                    'seconds': time.perf_counter() - t0}
        if args.dm_out and primary:
            numpy.save(args.dm_out, dm)
    if args.omega:
        with mol.with_range_coulomb(args.omega):
            opt = _vhf.VHFOpt(mol, direct_scf_tol=TOL)
    progress('%s: %d timed FP64 builds, per-class times' % (name, steps))
    r64 = gpu_leg(opt.basis, dm, 0.0, steps, warmup, class_reps, args.omega)
    rmx = None
    if fp32_tol > 0:
        progress('%s: %d timed mixed builds' % (name, steps))
        rmx = gpu_leg(opt.basis, dm, fp32_tol, steps, warmup, 0, args.omega)
    emx = None
    if rmx is not None and e64 is not None and not args.omega:
        progress('%s: mixed-precision RHF' % name)
        mfm = rhf.RHF(mol, direct_scf_tol=TOL, fp32_tol=fp32_tol)
        mfm.conv_tol = 1e-10
        emx = mfm.kernel()
    cpu = run_cpu_child(args, name, args.omega, dm, r64, rmx, with_oracle, with_cpu)
    return {'mol': mol, 'desc': desc, 'nao': nao, 'scf': scf_info, 'e64': e64, 'emx': emx,
            'r64': r64, 'rmx': rmx, 'accuracy': cpu.get('accuracy'), 'cpu': cpu.get('cpu_baseline')}


def main_single(args):
    res = single_molecule(args, args.workload, args.steps, args.warmup, not args.no_cpu, not args.no_oracle,
                          args.fp32_tol, args.class_reps)
    r64, rmx = res['r64'], res['rmx']
    t_total, lat, classes, st = r64['t_total'], r64['lat'], r64['classes'], r64['stats']
    builds = args.steps / t_total
    quartet_ms = st['ms_quartets']
    total_flops = st['flops']
    out = {
        'metric': METRIC, 'value': builds, 'unit': 'JK builds/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': t_total / args.steps * 1e3, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64',
        'data': 'synthetic geometry; converged RHF density from this engine',
        'config': {'workload': res['desc'] + ', one J/K build of the converged RHF density, hermi=1, '
                   'direct_scf_tol=1e-13', 'nao': res['nao'], 'n_dm': 1, 'precision': 'fp64',
                   'omega': args.omega},
        'quartets_per_s': st['quartets'] * builds, 'quartets_per_build': int(st['quartets']),
        'p50_ms': float(numpy.percentile(lat, 50) * 1e3), 'p99_ms': float(numpy.percentile(lat, 99) * 1e3),
        'scf': res['scf'],
        'counted_build': {'what': 'one build with statistics (eager, HIP events on the build stream)',
                          'ms_quartet_phase': quartet_ms, 'ms_transforms': st['ms_transform'],
                          'fp64_tflops_quartet_phase': total_flops / (quartet_ms * 1e-3) / 1e12,
                          'frac_of_fp64_peak': total_flops / (quartet_ms * 1e-3) / FP64_PEAK,
                          'flops': total_flops},
    }
    if classes:
        dom = classes[0]
        out['roofline'] = {
            'bound': 'valu', 'compute': 'FP64 vector ALU (no MFMA: Rys recurrences)',
            'kernel': dom['kernel'], 'class': dom['class'], 'achieved': dom['tflops'], 'peak': FP64_PEAK / 1e12,
            'unit': 'TFLOP/s', 'frac': dom['tflops'] * 1e12 / FP64_PEAK,
            'traffic': traffic_bytes(dom['kernel'], args.workload),
            'flops_per_launch': dom['flops'], 'ms_per_launch': dom['ms_per_launch'],
            'timing': 'the class alone on the build stream (class filter), HIP events around its launch, '
                      'mean of %d launches' % dom['launches']}
        out['classes'] = classes
        out['class_sum_ms'] = sum(c['ms_per_launch'] for c in classes)
    if rmx is not None:
        b_mx = args.steps / rmx['t_total']
        smx = rmx['stats']
        out['mixed'] = {
            'fp32_tol': args.fp32_tol, 'jk_builds_per_s': b_mx, 'speedup_vs_fp64': b_mx / builds,
            'quartets_per_s': smx['quartets'] * b_mx,
            'fp32_quartet_fraction': smx['quartets_fp32'] / max(smx['quartets'], 1),
            'p50_ms': float(numpy.percentile(rmx['lat'], 50) * 1e3),
            'p99_ms': float(numpy.percentile(rmx['lat'], 99) * 1e3),
            'max_abs_dJ_vs_fp64_gpu': float(abs(rmx['vj'] - r64['vj']).max()),
            'max_abs_dK_vs_fp64_gpu': float(abs(rmx['vk'] - r64['vk']).max()),
            'e_tot': res['emx'], 'dE_vs_fp64': res['emx'] - res['e64'] if res['emx'] is not None else None,
        }
    if res['accuracy'] is not None:
        out['accuracy'] = res['accuracy']
    if res['cpu'] is not None:
        out['cpu_baseline'] = res['cpu']
    if args.workload == 'c3' and not args.no_secondary and not args.dm_in and not args.omega:
        progress('secondary: (H2O)16 def2-SVP')
        sec = single_molecule(args, 'c2', args.steps, args.warmup, False, not args.no_oracle, 0.0, 0, primary=False)
        s64 = sec['r64']
        out['secondary'] = {'workload': sec['desc'], 'jk_builds_per_s': args.steps / s64['t_total'],
                            'ms_per_build': s64['t_total'] / args.steps * 1e3,
                            'p50_ms': float(numpy.percentile(s64['lat'], 50) * 1e3),
                            'quartets_per_build': int(s64['stats']['quartets']),
                            'scf': sec['scf'], 'accuracy': sec['accuracy']}
    print(json.dumps(out))


def bench_batched(args):
    """256 random-walk chains (def2-SVP): batched SCF (one graph-replayed build
    per iteration for all molecules still active) against the same molecules'
    own RHF runs, one molecule per build; both count active-molecule builds
    (each a density-difference build of one molecule)."""
    from pyscf_amd import gto, workloads
    from pyscf_amd.scf import batch, rhf
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
        n_act = sum(b.build_active)
        cyc = [{'active': a, 'ms': 1e3 * t} for a, t in zip(b.build_active, b.build_times)]
        res[label] = {'e_tot': e, 'scf_s': t_scf, 'iterations': b.cycles, 'jk_s': t_jk,
                      'converged': int(b.converged.sum()), 'conv_mask': numpy.array(b.converged),
                      'active_builds': n_act, 'active_builds_per_s': n_act / t_jk, 'per_build': cyc}
    progress('c4: unbatched baseline, each molecule its own FP64 RHF')
    n_single, t_single, e_single = 0, 0.0, []
    for m in mols:
        mf = rhf.RHF(m, direct_scf_tol=TOL)
        mf.conv_tol = 1e-9
        e_single.append(mf.kernel())
        n_single += len(mf.jk_times)
        t_single += sum(mf.jk_times)
    single = n_single / t_single
    e_single = numpy.array(e_single)
    de = numpy.abs(res['mixed']['e_tot'] - res['fp64']['e_tot'])
    both = res['mixed']['conv_mask'] & res['fp64']['conv_mask']
    de_single = numpy.abs(res['fp64']['e_tot'] - e_single)
    pb = res['mixed']['per_build']
    out = {
        'metric': METRIC, 'value': res['mixed']['active_builds_per_s'], 'unit': 'active-molecule JK builds/s',
        'n_gpus': 1, 'steps': res['mixed']['iterations'], 'warmup': 0,
        'ms_per_step': 1e3 * res['mixed']['jk_s'] / res['mixed']['iterations'], 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64+f32', 'data': 'synthetic random-walk chains',
        'config': {'workload': '%d random-walk chains (3-12 heavy atoms, seeds [4,i]), def2-SVP, '
                   '%d-%d AOs, batched RHF to 1e-9' % (len(mols), min(naos), max(naos))},
        'fp64_batched_active_builds_per_s': res['fp64']['active_builds_per_s'],
        'mixed_batched_active_builds_per_s': res['mixed']['active_builds_per_s'],
        'unbatched_fp64_builds_per_s': single,
        'unbatched_builds': n_single, 'unbatched_jk_s': t_single,
        'speedup_mixed_batched_vs_unbatched_fp64': res['mixed']['active_builds_per_s'] / single,
        'speedup_fp64_batched_vs_unbatched_fp64': res['fp64']['active_builds_per_s'] / single,
        'first_build': pb[0], 'last_builds': pb[-5:],
        'max_abs_dE_mixed_vs_fp64': float(de[both].max()) if both.any() else None,
        'molecules_dE_over_1e-8': int((de[both] > 1e-8).sum()),
        'max_abs_dE_batched_vs_single_fp64': float(de_single[res['fp64']['conv_mask']].max()),
        'converged': [res['fp64']['converged'], res['mixed']['converged']],
        'scf_seconds': [res['fp64']['scf_s'], res['mixed']['scf_s']],
    }
    print(json.dumps(out))


def bench_serving(args):
    """Poisson stream of get_jk and RHF requests (50-max AOs, lognormal;
    70% def2-SVP, 30% def2-TZVP; the first request at the largest size, in
    def2-TZVP) through the continuous-batching server."""
    from pyscf_amd import gto, workloads
    from pyscf_amd.server import JKServer, Request
    rng = numpy.random.default_rng(5)
    reqs = []
    for i in range(args.nreq):
        basis = 'def2-svp' if rng.random() < 0.7 else 'def2-tzvp'
        target = float(numpy.clip(rng.lognormal(numpy.log(200), 0.8), 50, args.max_ao))
        if i == 0:          # the stream always holds one request at the largest size
            basis, target = 'def2-tzvp', float(args.max_ao)
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


def pmc_run(args):
    """Two J/K builds of a saved density of the workload, then a 64 MiB device
    copy (kernel k_copy) whose byte counts calibrate the HBM counters: it moves
    exactly 64 MiB each way."""
    from pyscf_amd.scf import _vhf
    mol, _ = workload(args.workload)
    dm = numpy.ascontiguousarray(numpy.load(args.pmc_run, allow_pickle=False))
    assert dm.shape == (mol.nao_nr(), mol.nao_nr())
    opt = _vhf.VHFOpt(mol, direct_scf_tol=TOL)
    eng = opt.engine
    eng.check(eng.lib.mjk_ctx_set_graphs(eng.handle, 0), 'set_graphs')   # kernels appear by name
    for _ in range(2):
        opt.basis.get_jk(dm[None], 1, True, True, TOL)
    lib, ctx = eng.lib, eng.handle
    nb = 64 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    eng.check(lib.mjk_device_alloc(ctx, nb, ctypes.byref(a)), 'alloc')
    eng.check(lib.mjk_device_alloc(ctx, nb, ctypes.byref(b)), 'alloc')
    eng.check(lib.mjk_device_copy(ctx, b, a, nb), 'copy')
    eng.check(lib.mjk_synchronize(ctx), 'sync')
    print('pmc run done', flush=True)


def _pmc_table(d, counter):
    """{kernel: [sum of counter per dispatch]} from a rocprofv3 --pmc CSV dir."""
    import csv
    files = glob.glob(os.path.join(d, '**', '*counter_collection.csv'), recursive=True)
    if not files:
        raise SystemExit('no counter_collection.csv under %s' % d)
    per = {}
    for fn in files:
        with open(fn) as f:
            for row in csv.DictReader(f):
                if row.get('Counter_Name') != counter:
                    continue
                key = (row['Kernel_Name'].split('(')[0], row['Dispatch_Id'])
                per[key] = per.get(key, 0.0) + float(row['Counter_Value'])
    out = {}
    for (k, _), v in per.items():
        out.setdefault(k, []).append(v)
    return out


def pmc_summary(args):
    fetch = _pmc_table(args.pmc_summary[0], 'FETCH_SIZE')
    write = _pmc_table(args.pmc_summary[1], 'WRITE_SIZE')
    copy = [k for k in fetch if 'k_copy' in k][0]
    fs = (64 << 20) / (fetch[copy][0] * 1024)      # KiB per dispatch -> bytes, calibrated
    ws = (64 << 20) / (write[copy][0] * 1024)
    kernels = {}
    for k in fetch:
        if 'k_jk' not in k or k not in write:
            continue
        kernels[k] = {'launches': len(fetch[k]),
                      'fetch_bytes_per_launch': fs * 1024 * float(numpy.mean(fetch[k])),
                      'write_bytes_per_launch': ws * 1024 * float(numpy.mean(write[k]))}
    print(json.dumps({'workload': args.workload, 'fetch_scale': fs, 'write_scale': ws,
                      'calibration': 'k_copy moves 64 MiB each way; FETCH_SIZE read %.0f KiB (scale %.4f), '
                                     'WRITE_SIZE read %.0f KiB (scale %.4f)'
                                     % (fetch[copy][0], fs, write[copy][0], ws),
                      'kernels': kernels}, indent=1))


def bench_df(args):
    """Density-fitted J/K of C30H62 cc-pVDZ with cc-pVDZ-JKFIT (pyscf/df/df_jk.py:278-418):
    the fitting tensor built once on the device, then timed J/K builds of the
    converged DF-RHF density from its occupied orbitals (the SCF path) and of
    the density itself (the general path), inputs and outputs resident in HBM."""
    from pyscf_amd import df
    from pyscf_amd.scf import _vhf, rhf
    mol, desc = workload('c3')
    nao = mol.nao_nr()
    eng = _vhf.get_engine()
    L = eng.lib
    progress('df: fitting tensor')
    t0 = time.perf_counter()
    d = df.DF(mol).build()
    t_fit = time.perf_counter() - t0
    naux = d.get_naoaux()
    progress('df: %d AOs, %d fitting functions, tensor %.1f GB in %.2f s' % (nao, naux, 8e-9 * naux * nao * nao, t_fit))
    progress('df: DF-RHF')
    mf = rhf.RHF(mol, direct_scf_tol=TOL).density_fit(with_df=d)
    mf.conv_tol = 1e-10
    t0 = time.perf_counter()
    e_df = mf.kernel()
    t_scf = time.perf_counter() - t0
    progress('df: device-resident DF-RHF')
    mfd = rhf.DeviceRHF(mol, direct_scf_tol=TOL).density_fit(with_df=d)
    mfd.conv_tol = 1e-10
    t0 = time.perf_counter()
    e_dfd = mfd.kernel()
    t_scf_dev = time.perf_counter() - t0
    progress('df: exact RHF for the fitting error')
    mfx = rhf.RHF(mol, direct_scf_tol=TOL)
    mfx.conv_tol = 1e-10
    e_exact = mfx.kernel()
    occ = mf.mo_occ > 0
    orbo = numpy.ascontiguousarray(mf.mo_coeff[:, occ] * numpy.sqrt(mf.mo_occ[occ]))
    dm = numpy.ascontiguousarray(orbo @ orbo.T)
    nocc = orbo.shape[1]
    # device copies of the inputs, device outputs
    bufs = {}
    for name, n in (('dm', nao * nao), ('orbo', nao * nocc), ('vj', nao * nao), ('vk', nao * nao)):
        ptr = ctypes.c_void_p()
        eng.check(L.mjk_device_alloc(eng.handle, 8 * n, ctypes.byref(ptr)), 'mjk_device_alloc')
        bufs[name] = ptr
    for name, a in (('dm', dm), ('orbo', orbo)):
        eng.check(L.mjk_memcpy_h2d(eng.handle, bufs[name], a.ctypes.data_as(ctypes.c_void_p), a.nbytes), 'h2d')
    eng.check(L.mjk_synchronize(eng.handle), 'sync')

    def timed(orbital_path, steps, warmup):
        po = bufs['orbo'] if orbital_path else None
        for _ in range(warmup):
            eng.check(L.mjk_df_get_jk_device(eng.handle, d.handle, bufs['dm'], 1, po, nocc, 1, 1, bufs['vj'],
                                             bufs['vk']), 'mjk_df_get_jk_device')
        eng.check(L.mjk_synchronize(eng.handle), 'sync')
        lat = []
        t0 = time.perf_counter()
        for _ in range(steps):
            t1 = time.perf_counter()
            eng.check(L.mjk_df_get_jk_device(eng.handle, d.handle, bufs['dm'], 1, po, nocc, 1, 1, bufs['vj'],
                                             bufs['vk']), 'mjk_df_get_jk_device')
            eng.check(L.mjk_synchronize(eng.handle), 'sync')
            lat.append(time.perf_counter() - t1)
        return (time.perf_counter() - t0) / steps, numpy.array(lat)

    progress('df: %d timed builds from occupied orbitals' % args.steps)
    t_mo, lat_mo = timed(True, args.steps, args.warmup)
    nsteps_gen = max(3, args.steps // 4)
    progress('df: %d timed builds of the density' % nsteps_gen)
    t_gen, _ = timed(False, nsteps_gen, 1)
    vj, vk = numpy.empty((nao, nao)), numpy.empty((nao, nao))
    eng.check(L.mjk_memcpy_d2h(eng.handle, vj.ctypes.data_as(ctypes.c_void_p), bufs['vj'], vj.nbytes), 'd2h')
    eng.check(L.mjk_memcpy_d2h(eng.handle, vk.ctypes.data_as(ctypes.c_void_p), bufs['vk'], vk.nbytes), 'd2h')
    vj2, vk2 = df.get_jk(d, mf.make_rdm1(mf.mo_coeff, mf.mo_occ), 1)
    for ptr in bufs.values():
        L.mjk_device_free(eng.handle, ptr)
    nn = nao * nao
    flops_j = 2 * 2.0 * naux * nn
    # K = Y^T Y is symmetric: its algorithmic FLOPs count each unique element once
    flops_mo = flops_j + 2.0 * naux * nao * nocc * nao + 1.0 * nao * (nao + 1) * naux * nocc
    flops_gen = flops_j + 2 * 2.0 * naux * nao * nn
    out = {
        'metric': 'DF J/K builds/s (fitted three-centre tensor resident in HBM)', 'value': 1.0 / t_mo,
        'unit': 'JK builds/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': 1e3 * t_mo, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f64', 'data': 'synthetic geometry; converged DF-RHF density from this engine',
        'config': {'workload': desc + ', cc-pVDZ-JKFIT (%d fitting functions), J and K of the converged '
                   'DF-RHF density from its %d occupied orbitals' % (naux, nocc)},
        'p50_ms': float(numpy.percentile(lat_mo, 50) * 1e3), 'p99_ms': float(numpy.percentile(lat_mo, 99) * 1e3),
        'general_density_path': {'ms_per_build': 1e3 * t_gen, 'builds_per_s': 1.0 / t_gen,
                                 'tflops': flops_gen / t_gen * 1e-12},
        'fit': {'seconds': t_fit, 'naux': naux, 'tensor_gb': 8e-9 * naux * nn},
        'scf': {'e_df': e_df, 'e_exact': e_exact, 'fitting_error_eh': e_df - e_exact, 'cycles': mf.cycles,
                'seconds': t_scf, 'device_resident': {'e_df': e_dfd, 'cycles': mfd.cycles, 'seconds': t_scf_dev,
                                                      'dE_vs_host_loop': e_dfd - e_df}},
        'paths_agree': {'max_dJ': float(abs(vj - vj2).max()), 'max_dK': float(abs(vk - vk2).max())},
        'roofline': {'bound': 'mfma', 'achieved': flops_mo / t_mo * 1e-12, 'peak': FP64_PEAK * 1e-12,
                     'unit': 'TFLOP/s', 'frac': flops_mo / t_mo / FP64_PEAK, 'traffic': None,
                     'note': 'whole build (J two passes over the tensor; K a strided-batched rocBLAS FP64 GEMM and '
                             'the symmetric product as lower blocks), algorithmic FLOPs (the symmetric product '
                             'counted once per unique element) over wall time per synchronised build'},
        'cpu_baseline': None,
    }
    print(json.dumps(out))


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py runs on one GPU only: the J/K engine has no multi-GPU path; '
              'got --gpus %d' % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.cpu_child:
        return cpu_child(args)
    if args.pmc_summary:
        return pmc_summary(args)
    if args.pmc_run:
        return pmc_run(args)
    if args.workload == 'c4':
        return bench_batched(args)
    if args.workload == 'c5':
        return bench_serving(args)
    if args.workload == 'c3df':
        return bench_df(args)
    return main_single(args)


if __name__ == '__main__':
    main()
