"""Tuning tables change how the quartet work is cut and which kernel runs,
not the integrals: J and K built under tables that set every row-kernel
class to other bra primitive pairs per task, other task sizes or the
fallback kernel equal the untuned build (and the oracle at the FP64
contract).  Table changes rebuild the task lists, also of a basis whose
build was captured as a HIP graph."""
import numpy
import pytest

from pyscf_amd import gto, tune, workloads
from pyscf_amd.scf import _vhf

pytestmark = pytest.mark.gpu
TOL = 1e-10
SCF_TOL = 1e-13


def density(nao, seed):
    c = numpy.random.default_rng(seed).standard_normal((nao, max(2, nao // 4))) / 4
    return c @ c.T


def table_for(basis, **setting):
    out = []
    for k, l4, ctr, nblocks, kern in basis.classes():
        if nblocks > 0 and kern in (1, 3):
            e = dict(zip(tune.FIELDS[:9], tune.class_key(basis, k)))
            e.update({'kernel': 0, 'rprim': 0, 'scale': 0.0})
            if kern == 3 and setting.get('kernel') == 2:
                continue
            e.update(setting)
            out.append(e)
    return out


@pytest.mark.parametrize('name', ['water4', 'butane'])
def test_tuned_equals_untuned(oracle, name):
    mol = {'water4': lambda: gto.M(atom=workloads.water_cluster(4, seed=3), basis='def2-svp'),
           'butane': lambda: gto.M(atom=workloads.alkane(4), basis='cc-pvdz')}[name]()
    eng = _vhf.get_engine()
    opt = _vhf.VHFOpt(mol, direct_scf_tol=SCF_TOL)
    b = opt.basis
    dm = density(mol.nao_nr(), 41)[None]
    saved = tune.read_table() if tune.load(eng) else []
    try:
        tune.set_table(eng, [])
        vj, vk = b.get_jk(dm, 1, True, True, SCF_TOL)
        for _ in range(2):            # the second repeat is captured as a graph
            b.get_jk(dm, 1, True, True, SCF_TOL)
        rows = [c for c in b.classes() if c[3] > 0 and c[4] in (1, 3)]
        assert rows
        for setting in ({'rprim': 1}, {'rprim': 16, 'scale': 0.25}, {'scale': 4.0}, {'kernel': 2}):
            table = table_for(b, **setting)
            tune.set_table(eng, table)
            for rep in range(3):      # eager, captured, replayed
                tj, tk = b.get_jk(dm, 1, True, True, SCF_TOL)
                dj, dk = abs(tj - vj).max(), abs(tk - vk).max()
                print('%s %s rep %d: max|dJ|=%.2e max|dK|=%.2e' % (name, setting, rep, dj, dk), flush=True)
                assert dj <= 1e-11 and dk <= 1e-11
            if setting.get('kernel') == 2:
                assert 1 not in [c[4] for c in b.classes() if c[3] > 0]
        if name == 'water4':
            j0, k0, _ = oracle.direct_full(mol, dm, vhfopt=oracle.VHFOpt(mol, SCF_TOL))
            k0 = oracle.hermi_triu(k0.copy(), 1)
            assert abs(vj - j0).max() <= TOL and abs(vk - k0).max() <= TOL
    finally:
        tune.set_table(eng, saved)


def test_context_loads_committed_table():
    eng = _vhf.get_engine()
    n = eng.lib.mjk_ctx_tuning_entries(eng.handle)
    import os
    if os.path.exists(tune.TABLE) and os.environ.get('MJK_TUNE', '1') != '0':
        assert n == len(tune.read_table())
    else:
        assert n == 0


def test_bad_tables():
    eng = _vhf.get_engine()
    saved = tune.read_table() if tune.load(eng) else []
    bad = {'la': 4, 'lb': 0, 'lc': 0, 'ld': 0, 'ca': 1, 'cb': 1, 'cc': 1, 'cd': 1, 'bucket': 3, 'kernel': 0, 'rprim': 1}
    try:
        with pytest.raises(RuntimeError, match='bad class key'):
            tune.set_table(eng, [bad])
        bad.update(la=0, kernel=3)
        with pytest.raises(RuntimeError, match='bad setting'):
            tune.set_table(eng, [bad])
    finally:
        tune.set_table(eng, saved)
