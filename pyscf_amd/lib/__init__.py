"""Loader for the HIP library and small numpy helpers.

``load_library`` mirrors pyscf/lib/misc.py:95-107 (ctypes load of an
in-tree shared object); there is no fallback: a missing or unloadable
library raises.  ``fp`` is pyscf/lib/misc.py:1259-1263 and ``hermi_triu``
pyscf/lib/numpy_helper.py:480-525.
"""

import ctypes
import os

import numpy

_LIBDIR = os.path.dirname(os.path.abspath(__file__))
_loaded = {}


class mjk_rhf_result(ctypes.Structure):
    _fields_ = [('e_tot', ctypes.c_double), ('gnorm', ctypes.c_double), ('cycles', ctypes.c_int),
                ('converged', ctypes.c_int)]


class mjk_request(ctypes.Structure):
    _fields_ = [('kind', ctypes.c_int), ('atm', ctypes.POINTER(ctypes.c_int)), ('natm', ctypes.c_int),
                ('bas', ctypes.POINTER(ctypes.c_int)), ('nbas', ctypes.c_int),
                ('env', ctypes.POINTER(ctypes.c_double)), ('env_size', ctypes.c_int),
                ('direct_scf_tol', ctypes.c_double), ('fp32_tol', ctypes.c_double),
                ('dm', ctypes.POINTER(ctypes.c_double)), ('n_dm', ctypes.c_int), ('hermi', ctypes.c_int),
                ('with_j', ctypes.c_int), ('with_k', ctypes.c_int), ('nelectron', ctypes.c_int),
                ('max_cycle', ctypes.c_int), ('conv_tol', ctypes.c_double)]


class mjk_request_info(ctypes.Structure):
    _fields_ = [('state', ctypes.c_int), ('nao', ctypes.c_int), ('n_dm', ctypes.c_int), ('cycles', ctypes.c_int),
                ('converged', ctypes.c_int), ('slot', ctypes.c_int), ('e_tot', ctypes.c_double),
                ('t_submit', ctypes.c_double), ('t_start', ctypes.c_double), ('t_done', ctypes.c_double),
                ('hbm_estimate', ctypes.c_size_t), ('hbm_peak', ctypes.c_size_t), ('error', ctypes.c_char * 256)]


class mjk_stats(ctypes.Structure):
    _fields_ = [('ms_total', ctypes.c_double),
                ('ms_quartets', ctypes.c_double),
                ('ms_transform', ctypes.c_double),
                ('quartets', ctypes.c_longlong),
                ('prim_quartets', ctypes.c_longlong),
                ('flops', ctypes.c_double),
                ('quartets_fp32', ctypes.c_longlong),
                ('flops_fp32', ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _declare(lib):
    c_int, c_double, c_void_p, c_size_t = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t
    P = ctypes.POINTER
    lib.mjk_ctx_create.argtypes = [P(c_void_p), c_int, c_size_t]
    lib.mjk_ctx_destroy.argtypes = [c_void_p]
    lib.mjk_ctx_destroy.restype = None
    lib.mjk_last_error.argtypes = [c_void_p]
    lib.mjk_last_error.restype = ctypes.c_char_p
    lib.mjk_ctx_hbm_used.argtypes = [c_void_p]
    lib.mjk_ctx_hbm_used.restype = c_size_t
    lib.mjk_ctx_set_row_lds.argtypes = [c_void_p, c_size_t]
    lib.mjk_ctx_set_graphs.argtypes = [c_void_p, c_int]
    lib.mjk_basis_upload.argtypes = [c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_double, P(c_void_p)]
    lib.mjk_basis_upload_batch.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_double, P(c_void_p)]
    lib.mjk_basis_destroy.argtypes = [c_void_p]
    lib.mjk_basis_destroy.restype = None
    lib.mjk_basis_nao.argtypes = [c_void_p, c_int]
    lib.mjk_basis_set_active.argtypes = [c_void_p, c_void_p, c_void_p]
    lib.mjk_basis_slot_size.argtypes = [c_void_p]
    lib.mjk_basis_slot_size.restype = ctypes.c_longlong
    lib.mjk_basis_q_cond.argtypes = [c_void_p, c_void_p, c_int, c_void_p]
    lib.mjk_get_jk.argtypes = [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_double, c_double,
                               c_void_p, c_void_p, P(mjk_stats)]
    lib.mjk_get_jk_device.argtypes = lib.mjk_get_jk.argtypes
    lib.mjk_get_jk_device_noalloc.argtypes = lib.mjk_get_jk.argtypes[:-1]
    lib.mjk_get_jk_resident.argtypes = [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_double, c_double,
                                        P(mjk_stats)]
    lib.mjk_basis_fetch_jk.argtypes = [c_void_p, c_void_p, c_int, c_void_p, c_void_p]
    lib.mjk_basis_set_class_filter.argtypes = [c_void_p, c_void_p, c_int]
    lib.mjk_basis_nclasses.argtypes = [c_void_p]
    lib.mjk_basis_class_key.argtypes = [c_void_p, c_void_p, c_int, c_void_p]
    lib.mjk_ctx_set_tuning.argtypes = [c_void_p, c_int, c_void_p, c_void_p]
    lib.mjk_ctx_tuning_entries.argtypes = [c_void_p]
    lib.mjk_rhf_create.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_double, c_int, P(c_void_p)]
    lib.mjk_rhf_run.argtypes = [c_void_p, c_void_p, c_void_p, c_int, c_double, c_double, c_double, c_double,
                                c_double, P(mjk_rhf_result)]
    lib.mjk_rhf_fetch.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
    lib.mjk_rhf_set_df.argtypes = [c_void_p, c_void_p, c_void_p]
    lib.mjk_rhf_device_density.argtypes = [c_void_p]
    lib.mjk_rhf_device_density.restype = c_void_p
    lib.mjk_rhf_destroy.argtypes = [c_void_p]
    lib.mjk_rhf_destroy.restype = None
    lib.mjk_int1e_device.argtypes = [c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_int]
    lib.mjk_df_build.argtypes = [c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p,
                                 c_int, c_void_p, c_double, P(c_void_p)]
    lib.mjk_df_naux.argtypes = [c_void_p]
    lib.mjk_df_get_jk.argtypes = [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int, c_int, c_int, c_void_p,
                                  c_void_p]
    lib.mjk_df_get_jk_device.argtypes = lib.mjk_df_get_jk.argtypes
    lib.mjk_df_fetch.argtypes = [c_void_p, c_void_p, c_int, c_void_p]
    lib.mjk_df_destroy.argtypes = [c_void_p]
    lib.mjk_df_destroy.restype = None
    lib.mjk_server_create.argtypes = [c_int, c_int, c_size_t, P(c_void_p)]
    lib.mjk_submit.argtypes = [c_void_p, P(mjk_request), P(ctypes.c_longlong)]
    lib.mjk_poll.argtypes = [c_void_p, ctypes.c_longlong, P(mjk_request_info)]
    lib.mjk_wait_any.argtypes = [c_void_p, c_double, P(ctypes.c_longlong)]
    lib.mjk_result.argtypes = [c_void_p, ctypes.c_longlong, c_void_p, c_void_p, c_void_p, c_void_p]
    lib.mjk_server_hbm_held.argtypes = [c_void_p]
    lib.mjk_server_hbm_held.restype = c_size_t
    lib.mjk_server_error.argtypes = [c_void_p]
    lib.mjk_server_error.restype = ctypes.c_char_p
    lib.mjk_server_destroy.argtypes = [c_void_p]
    lib.mjk_server_destroy.restype = None
    lib.mjk_basis_class.argtypes = [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p,
                                    P(ctypes.c_longlong), P(c_int)]
    lib.mjk_device_alloc.argtypes = [c_void_p, c_size_t, P(c_void_p)]
    lib.mjk_device_free.argtypes = [c_void_p, c_void_p]
    lib.mjk_memcpy_h2d.argtypes = [c_void_p, c_void_p, c_void_p, c_size_t]
    lib.mjk_memcpy_d2h.argtypes = [c_void_p, c_void_p, c_void_p, c_size_t]
    lib.mjk_synchronize.argtypes = [c_void_p]
    lib.mjk_mem_info.argtypes = [c_int, P(c_size_t), P(c_size_t)]
    lib.mjk_device_copy.argtypes = [c_void_p, c_void_p, c_void_p, c_size_t]
    lib.mjk_rys_roots.argtypes = [c_int, c_double, c_void_p, c_void_p]
    lib.mjk_c2s.argtypes = [c_int, c_void_p]
    lib.mjk_profile.argtypes = [c_void_p, c_void_p, c_int]
    lib.mjk_profile_read.argtypes = [c_void_p, c_void_p, c_int] + [c_void_p] * 6
    lib.mjk_profile_kinds.argtypes = [c_void_p, c_void_p, c_int, c_void_p]
    lib.mjk_profile_span.argtypes = [c_void_p, c_void_p, P(c_double), P(ctypes.c_longlong)]
    for name in ('mjk_int1e',):
        if hasattr(lib, name):
            getattr(lib, name).argtypes = [c_int, c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p]
    return lib


def library_path(name='libmjk'):
    return os.path.join(_LIBDIR, name + '.so')


def load_library(name='libmjk'):
    """ctypes handle of an in-tree library; raises if it is not built."""
    if name in _loaded:
        return _loaded[name]
    path = library_path(name)
    if not os.path.exists(path):
        raise OSError('%s is not built; run "python -m pyscf_amd.build"' % path)
    lib = _declare(ctypes.CDLL(path))
    _loaded[name] = lib
    return lib


def fp(a):
    """Fingerprint cos(arange(n)) . a.ravel() (pyscf/lib/misc.py:1259-1263)."""
    a = numpy.asarray(a)
    return numpy.dot(numpy.cos(numpy.arange(a.size)), a.ravel())


finger = fingerprint = fp


def hermi_triu(mat, hermi=1, inplace=True):
    """Fill the upper triangle from the lower one: hermi=1 mirror,
    hermi=2 anti-mirror (pyscf/lib/numpy_helper.py:480-525)."""
    if not inplace:
        mat = mat.copy()
    n = mat.shape[-1]
    iu = numpy.triu_indices(n, 1)
    if hermi == 1:
        mat[..., iu[0], iu[1]] = mat[..., iu[1], iu[0]].conj()
    elif hermi == 2:
        mat[..., iu[0], iu[1]] = -mat[..., iu[1], iu[0]].conj()
    else:
        raise ValueError('hermi must be 1 or 2')
    return mat


class _TaggedArray(numpy.ndarray):
    """ndarray carrying extra attributes (mo_coeff, mo_occ) through views."""

    def __array_finalize__(self, obj):
        if obj is not None:
            self.__dict__.update(getattr(obj, '__dict__', {}))


def tag_array(a, **kwargs):
    """a with attributes attached, as pyscf.lib.tag_array
    (pyscf/lib/numpy_helper.py): the density-fitted K builds use a density's
    mo_coeff/mo_occ when they are there."""
    t = numpy.asarray(a).view(_TaggedArray)
    t.__dict__.update(kwargs)
    return t
