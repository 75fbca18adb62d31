"""C-ABI library: loads, exports every declared symbol, validates args
without a GPU, and its Rys table reproduces the Boys moments."""
import ctypes
import os
import re

import numpy
import pytest

from pyscf_amd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, 'include', 'mjk.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'\b(mjk_[a-z0-9_]+)\s*\(', src)))


def test_library_exports_declared_symbols():
    l = lib.load_library()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(l, n), n


def test_bad_arguments_return_codes():
    l = lib.load_library()
    x = numpy.zeros(3)
    w = numpy.zeros(3)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert l.mjk_rys_roots(0, 1.0, p(x), p(w)) == -1
    assert l.mjk_rys_roots(10, 1.0, p(x), p(w)) == -1
    assert l.mjk_rys_roots(3, -1.0, p(x), p(w)) == -1
    assert l.mjk_ctx_create(None, 0, 0) == -1
    assert l.mjk_last_error(None) == b'no context'
    assert l.mjk_ctx_set_row_lds(None, 0) == -1
    assert l.mjk_ctx_set_graphs(None, 1) == -1


@pytest.mark.parametrize('n', range(1, 10))
def test_rys_table_matches_boys_moments(oracle, n):
    l = lib.load_library()
    rng = numpy.random.default_rng(n)
    Ts = numpy.concatenate([[0.0, 1e-10, 2.0, 36.0 + 6 * n - 1e-9, 36.0 + 6 * n, 500.0],
                            rng.uniform(0, 40 + 6 * n, 40)])
    for T in Ts:
        x = numpy.zeros(n)
        w = numpy.zeros(n)
        assert l.mjk_rys_roots(n, float(T), x.ctypes.data_as(ctypes.c_void_p), w.ctypes.data_as(ctypes.c_void_p)) == 0
        for m in range(2 * n):
            ref = float(oracle.boys_mp(m, T))
            assert abs(numpy.dot(w, x**m) - ref) <= 2e-14 * ref, (n, T, m)
        xo, wo = oracle.rys_roots(n, float(T))
        assert abs(x - xo).max() < 1e-13 and (abs(w - wo) / wo).max() < 1e-12


def test_server_and_scf_bad_arguments():
    l = lib.load_library()
    h = ctypes.c_void_p()
    t = ctypes.c_longlong()
    assert l.mjk_server_create(0, 0, 0, ctypes.byref(h)) == -1
    assert l.mjk_server_create(0, 65, 0, ctypes.byref(h)) == -1
    assert l.mjk_submit(None, None, ctypes.byref(t)) == -1
    assert l.mjk_poll(None, 1, None) == -1
    assert l.mjk_wait_any(None, 0.0, ctypes.byref(t)) == -1
    assert l.mjk_result(None, 1, None, None, None, None) == -1
    assert l.mjk_server_hbm_held(None) == 0
    l.mjk_server_destroy(None)
    assert l.mjk_rhf_create(None, None, None, None, 1, 0.0, 8, ctypes.byref(h)) == -1
    assert l.mjk_rhf_run(None, None, None, 10, 1e-9, 1e-5, 1e-13, 0.0, 0.0, None) == -1
    assert l.mjk_rhf_fetch(None, None, None, None, None, None) == -1
    assert l.mjk_rhf_device_density(None) is None
    l.mjk_rhf_destroy(None)
    assert l.mjk_ctx_tuning_entries(None) == -1
