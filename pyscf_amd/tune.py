"""Per-class autotuning of the J/K build.

The engine's task lists have three per-class knobs whose best values depend
on the class and its size: the kernel (row kernel or its fallback: the tile
kernel up to 81 Cartesian components, the cooperative kernel above; for a
merged class, its split pairs), the bra primitive pairs per row task and the
size of the row tasks.  The library's default rules are fixed
(``build_tiles`` in ``csrc/mjk_api.hip``); a tuning table overrides them per
key (la lb lc ld, contraction pattern, work bucket = floor(log2) of the
class's Schwarz-prescreened primitive quartets), the nearest bucket within 2
applying (``mjk_ctx_set_tuning``).

* The table is a text file, one entry a line: ``la lb lc ld ca cb cc cd
  bucket kernel rprim scale`` and then, for the reader, the class's default
  and tuned times and where they were measured; ``#`` starts a comment.
  ``mjk_ctx_create`` loads ``TABLE`` (or the file ``MJK_TUNING`` names)
  unless ``MJK_TUNE=0``; ``load(engine, path)`` installs another.
* ``python -m pyscf_amd.tune [--out PATH]`` measures one: for every class of
  a set of molecules, each class alone (builds restricted to it, HIP events
  around its launch, mean of the timed builds) under candidate settings by
  coordinate descent -- bra primitive pairs per task, then the task-size
  scale, then the kernel -- keeping a setting only when it beats the current
  one by more than 3%.  Results agree with untuned builds to rounding: the
  knobs change how the work is cut and which kernel runs, not the integrals.
"""

import argparse
import ctypes
import os
import sys
import time

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, 'tuning', 'mi355x.txt')
FIELDS = ('la', 'lb', 'lc', 'ld', 'ca', 'cb', 'cc', 'cd', 'bucket', 'kernel', 'rprim')


def table_arrays(entries):
    """(n, int32 keys [n, 11], float64 scale [n]) of a list of entry dicts."""
    keys = numpy.array([[int(e[f]) for f in FIELDS] for e in entries], dtype=numpy.int32).reshape(-1, 11)
    scale = numpy.array([float(e.get('scale', 0.0)) for e in entries], dtype=numpy.double)
    return len(entries), numpy.ascontiguousarray(keys), numpy.ascontiguousarray(scale)


def read_table(path=TABLE):
    """Entries (dicts) of a table file."""
    out = []
    with open(path) as f:
        for line in f:
            line = line.split('#', 1)[0].split()
            if not line:
                continue
            e = {k: int(v) for k, v in zip(FIELDS, line[:11])}
            e['scale'] = float(line[11])
            if len(line) >= 14:
                e['ms_default'], e['ms_tuned'] = float(line[12]), float(line[13])
                e['from'] = ' '.join(line[14:])
            out.append(e)
    return out


def write_table(path, entries, header=()):
    entries = sorted(entries, key=lambda e: tuple(e[f] for f in FIELDS[:9]))
    with open(path, 'w') as f:
        for h in header:
            f.write('# %s\n' % h)
        f.write('# la lb lc ld ca cb cc cd bucket kernel rprim scale ms_default ms_tuned measured_on\n')
        for e in entries:
            f.write('%s %g %.4f %.4f %s\n' % (' '.join(str(int(e[k])) for k in FIELDS), float(e.get('scale', 0.0)),
                                               e.get('ms_default', 0.0), e.get('ms_tuned', 0.0), e.get('from', '')))


def set_table(engine, entries):
    n, keys, scale = table_arrays(entries)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    engine.check(engine.lib.mjk_ctx_set_tuning(engine.handle, n, p(keys), p(scale)), 'mjk_ctx_set_tuning')


def load(engine, path=TABLE):
    """Install the table at `path` on the engine's context; False if absent."""
    if not os.path.exists(path):
        return False
    set_table(engine, read_table(path))
    return True


def class_key(basis, k):
    key = numpy.zeros(9, dtype=numpy.int32)
    eng = basis.engine
    eng.check(eng.lib.mjk_basis_class_key(eng.handle, basis.handle, k, key.ctypes.data_as(ctypes.c_void_p)),
              'mjk_basis_class_key')
    return tuple(int(x) for x in key)


def time_class(basis, k, tol, reps):
    """Mean GPU time (ms) of class k alone over `reps` statistics builds (one
    untimed build first, which also rebuilds task lists after a table change)."""
    basis.set_class_filter(k)
    try:
        basis.get_jk_resident(1, 1, True, True, tol, stats=True)
        ms = [basis.get_jk_resident(1, 1, True, True, tol, stats=True)['ms_quartets'] for _ in range(reps)]
    finally:
        basis.set_class_filter(-1)
    return float(numpy.mean(ms))


def tune_molecule(engine, b, label, dm, table, tol=1e-13, reps=8, log=print):
    """Coordinate descent over the classes of basis handle b (one molecule,
    density dm); updates `table` (dict key -> entry) in place."""
    b.get_jk(dm[None], 1, True, True, tol)          # resident density
    for k, l4, ctr, nblocks, kern in b.classes():
        if nblocks == 0 or kern not in (1, 3):      # the knobs act on row-kernel classes
            continue
        key = class_key(b, k)
        base = dict(table)
        cur = dict(table.get(key, {'kernel': 0, 'rprim': 0, 'scale': 0.0}))
        set_table(engine, list(base.values()))
        t_cur = time_class(b, k, tol, reps)
        t_default = t_cur

        def trial(setting):
            e = dict(zip(FIELDS[:9], key))
            e.update(setting)
            trial_table = dict(base)
            trial_table[key] = e
            set_table(engine, list(trial_table.values()))
            return time_class(b, k, tol, reps)

        steps = [('rprim', v) for v in (1, 2, 4, 8, 16)] + [('scale', v) for v in (0.25, 4.0)]
        if kern == 1:
            steps.append(('kernel', 2))
        for field, value in steps:
            if cur.get(field) == value:
                continue
            cand = dict(cur)
            cand[field] = value
            t = trial(cand)
            if t < 0.97 * t_cur:
                cur, t_cur = cand, t
        if t_cur < 0.97 * t_default:
            e = dict(zip(FIELDS[:9], key))
            e.update(cur)
            e['ms_tuned'] = round(t_cur, 4)
            e['ms_default'] = round(t_default, 4)
            e['from'] = label
            table[key] = e
        log('class %-10s ctr %s bucket %2d: default %.3f ms, tuned %.3f ms %s'
            % ('(%d%d|%d%d)' % l4, ''.join(map(str, ctr)), key[8], t_default, t_cur,
               {f: cur[f] for f in ('kernel', 'rprim', 'scale') if f in cur}))
    set_table(engine, list(table.values()))


def tuning_molecules():
    """(name, mol) of the sweep: the single-molecule benchmark systems and a
    def2-TZVP chain with f shells."""
    from pyscf_amd import gto, workloads
    return [('H2O cc-pVDZ', gto.M(atom=workloads.h2o(), basis='cc-pvdz')),
            ('(H2O)16 def2-SVP', gto.M(atom=workloads.water_cluster(16, seed=1), basis='def2-svp')),
            ('C30H62 cc-pVDZ', gto.M(atom=workloads.alkane(30), basis='cc-pvdz')),
            ('chain def2-TZVP', gto.M(atom=workloads.random_chain([9, 0], nmin=12, nmax=12), basis='def2-tzvp'))]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--out', default=TABLE)
    ap.add_argument('--reps', type=int, default=8)
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    from pyscf_amd.scf import _vhf, rhf
    engine = _vhf.get_engine()
    set_table(engine, [])                           # measure against the library's default rules
    table = {}
    t0 = time.perf_counter()
    for name, mol in tuning_molecules():
        print('[tune %.0fs] %s (%d AOs): RHF density' % (time.perf_counter() - t0, name, mol.nao_nr()), flush=True)
        mf = rhf.RHF(mol, direct_scf_tol=1e-13)
        mf.conv_tol = 1e-8
        mf.max_cycle = 40
        mf.kernel()
        tune_molecule(engine, mf._opt[None].basis, '%s (%d AOs)' % (name, mol.nao_nr()),
                      numpy.ascontiguousarray(mf.dm), table, reps=args.reps,
                      log=lambda s: print('[tune %.0fs] %s' % (time.perf_counter() - t0, s), flush=True))
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    write_table(args.out, table.values(),
                header=('Per-class J/K tuning table, MI355X (gfx950); written by python -m pyscf_amd.tune',
                        'kernel: 0 default, 1 row kernel, 2 fallback; rprim: bra primitive pairs per row task '
                        '(0 default); scale: row task size factor (0 default); times: the class alone, ms'))
    print('[tune %.0fs] %d entries written to %s' % (time.perf_counter() - t0, len(table), args.out), flush=True)


if __name__ == '__main__':
    main()
