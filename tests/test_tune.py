"""Tuning table format (CPU): the committed table's entries, and the
conversion to the arrays mjk_ctx_set_tuning takes."""
import os

import numpy
import pytest

from pyscf_amd import tune


def test_table_arrays():
    entries = [{'la': 1, 'lb': 0, 'lc': 1, 'ld': 0, 'ca': 1, 'cb': 1, 'cc': 1, 'cd': 1, 'bucket': 17,
                'kernel': 0, 'rprim': 4, 'scale': 0.25},
               {'la': 0, 'lb': 0, 'lc': 0, 'ld': 0, 'ca': 2, 'cb': 1, 'cc': 2, 'cd': 1, 'bucket': 9,
                'kernel': 2, 'rprim': 0}]
    n, keys, scale = tune.table_arrays(entries)
    assert n == 2 and keys.dtype == numpy.int32 and keys.shape == (2, 11) and keys.flags.c_contiguous
    assert keys[0].tolist() == [1, 0, 1, 0, 1, 1, 1, 1, 17, 0, 4]
    assert keys[1].tolist() == [0, 0, 0, 0, 2, 1, 2, 1, 9, 2, 0]
    assert scale.tolist() == [0.25, 0.0]
    n, keys, scale = tune.table_arrays([])
    assert n == 0 and keys.shape == (0, 11) and scale.shape == (0,)


def test_table_file_round_trip(tmp_path):
    entries = [{'la': 2, 'lb': 1, 'lc': 1, 'ld': 1, 'ca': 1, 'cb': 1, 'cc': 1, 'cd': 1, 'bucket': 20,
                'kernel': 1, 'rprim': 2, 'scale': 4.0, 'ms_default': 1.5, 'ms_tuned': 1.2, 'from': 'C30H62 (730 AOs)'}]
    path = str(tmp_path / 't.txt')
    tune.write_table(path, entries, header=('test',))
    back = tune.read_table(path)
    assert back == entries


@pytest.mark.skipif(not os.path.exists(tune.TABLE), reason='no tuning table committed')
def test_committed_table():
    entries = tune.read_table(tune.TABLE)
    assert entries
    keys = set()
    for e in entries:
        assert all(0 <= e[f] <= 3 for f in ('la', 'lb', 'lc', 'ld'))
        assert e['la'] >= e['lb'] and e['lc'] >= e['ld']
        assert all(e[f] in (1, 2) for f in ('ca', 'cb', 'cc', 'cd'))
        assert e['kernel'] in (0, 1, 2) and 0 <= e['rprim'] <= 4096 and 0 <= e['scale'] <= 1e6
        # an entry is kept only for a measured gain
        assert e['ms_tuned'] < e['ms_default']
        key = tuple(e[f] for f in tune.FIELDS[:9])
        assert key not in keys
        keys.add(key)
    tune.table_arrays(entries)
