"""Density-fitted J/K and SCF (pyscf/df/df_jk.py).

``get_jk(dfobj, dm, hermi, with_j, with_k)`` follows pyscf/df/df_jk.py:278-418:
J = sum_P B_P tr(B_P D) for every density; K = sum_P B_P D B_P, through the
occupied orbitals X = C_occ sqrt(n_occ) (K = sum_P (B_P X)(B_P X)^T) when the
density carries mo_coeff and mo_occ (the SCF's densities do).  Complex
densities are split into real and imaginary parts.

``density_fit(mf, auxbasis)`` returns the SCF object with its J/K replaced
(pyscf/df/df_jk.py:32-125): full (non-incremental) J/K every cycle, as the
reference turns direct_scf off for density fitting.
"""

import ctypes

import numpy

from pyscf_amd import lib as _lib
from pyscf_amd.df.df import DF


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _jk_real(dfobj, dms, orbo, with_j, with_k):
    eng = dfobj.engine
    n_dm, nao = dms.shape[0], dms.shape[-1]
    vj = numpy.empty_like(dms) if with_j else None
    vk = numpy.empty_like(dms) if with_k else None
    nocc = 0 if orbo is None else orbo.shape[-1]
    eng.check(eng.lib.mjk_df_get_jk(eng.handle, dfobj.handle, _p(dms), n_dm, _p(orbo), nocc, int(bool(with_j)),
                                    int(bool(with_k)), _p(vj), _p(vk)), 'mjk_df_get_jk')
    return vj, vk


def get_jk(dfobj, dm, hermi=1, with_j=True, with_k=True, direct_scf_tol=1e-13):
    assert with_j or with_k
    dfobj.build()
    dms = numpy.asarray(dm)
    shape = dms.shape
    nao = shape[-1]
    dms = dms.reshape(-1, nao, nao)
    if numpy.iscomplexobj(dms):
        re = numpy.ascontiguousarray(dms.real)
        im = numpy.ascontiguousarray(dms.imag)
        jr, kr = _jk_real(dfobj, re, None, with_j, with_k)
        ji, ki = _jk_real(dfobj, im, None, with_j, with_k)
        vj = (jr + 1j * ji).reshape(shape) if with_j else None
        vk = (kr + 1j * ki).reshape(shape) if with_k else None
        return vj, vk
    dms = numpy.ascontiguousarray(dms, dtype=numpy.double)
    orbo = None
    mo_coeff = getattr(dm, 'mo_coeff', None)
    if with_k and mo_coeff is not None:
        mo_coeff = numpy.asarray(mo_coeff).reshape(-1, nao, numpy.asarray(mo_coeff).shape[-1])
        mo_occ = numpy.asarray(dm.mo_occ).reshape(-1, mo_coeff.shape[-1])
        if mo_coeff.shape[0] == dms.shape[0]:
            nocc = max(int((o > 0).sum()) for o in mo_occ)
            orbo = numpy.zeros((dms.shape[0], nao, nocc))
            for k in range(dms.shape[0]):
                occ = mo_occ[k] > 0
                x = mo_coeff[k][:, occ] * numpy.sqrt(mo_occ[k][occ])
                orbo[k, :, :x.shape[1]] = x
    vj, vk = _jk_real(dfobj, dms, orbo, with_j, with_k)
    return (vj.reshape(shape) if with_j else None), (vk.reshape(shape) if with_k else None)


def density_fit(mf, auxbasis=None, with_df=None):
    """mf with density-fitted J/K (the reference's mf.density_fit())."""
    from pyscf_amd.scf import rhf

    base = type(mf)

    class DFHF(base):
        def get_jk(self, mol=None, dm=None, hermi=1, with_j=True, with_k=True, omega=None):
            t0 = _lib_time()
            vj, vk = self.with_df.get_jk(dm, hermi, with_j, with_k, omega=omega)
            self.jk_times.append(_lib_time() - t0)
            return vj, vk

        def get_veff(self, mol=None, dm=None, dm_last=None, vhf_last=None, hermi=1):
            vj, vk = self.get_jk(mol, dm, hermi)
            return vj - vk * .5

        def make_rdm1(self, mo_coeff, mo_occ):
            return _lib.tag_array(base.make_rdm1(mo_coeff, mo_occ), mo_coeff=mo_coeff, mo_occ=mo_occ)

    if issubclass(base, rhf.DeviceRHF):
        # the device-resident loop takes the fit itself (mjk_rhf_set_df)
        new = base.__new__(base)
        new.__dict__.update(mf.__dict__)
        new.with_df = with_df or DF(mf.mol, auxbasis)
        return new
    new = DFHF.__new__(DFHF)
    new.__dict__.update(mf.__dict__)
    new.with_df = with_df or DF(mf.mol, auxbasis)
    return new


def _lib_time():
    import time
    return time.perf_counter()
