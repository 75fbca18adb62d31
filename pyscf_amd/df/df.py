"""The density-fitting object (counterpart of pyscf.df.DF, pyscf/df/df.py).

``build()`` computes (P|Q), (ij|P) and the fitted tensor B = L^-1 (P|ij) on
the device (``mjk_df_build``); B stays in HBM for the object's life."""

import ctypes

import numpy

from pyscf_amd.df.addons import make_auxmol

LINEAR_DEP_THR = 1e-7   # pyscf/df/incore.py:33


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class DF:
    def __init__(self, mol, auxbasis=None, engine=None):
        self.mol = mol
        self.auxbasis = auxbasis
        self.auxmol = None
        self.engine = engine
        self.lindep = LINEAR_DEP_THR
        self.handle = None

    def build(self):
        from pyscf_amd.scf import _vhf
        self.engine = self.engine or _vhf.get_engine()
        if self.handle is not None:
            return self
        self.auxmol = make_auxmol(self.mol, self.auxbasis)
        eng, L = self.engine, self.engine.lib
        a = [numpy.ascontiguousarray(x, dtype=t) for x, t in
             ((self.mol._atm, numpy.int32), (self.mol._bas, numpy.int32), (self.mol._env, numpy.double),
              (self.auxmol._atm, numpy.int32), (self.auxmol._bas, numpy.int32), (self.auxmol._env, numpy.double))]
        h = ctypes.c_void_p()
        eng.check(L.mjk_df_build(eng.handle, _p(a[0]), len(a[0]), _p(a[1]), len(a[1]), _p(a[2]),
                                 _p(a[3]), len(a[3]), _p(a[4]), len(a[4]), _p(a[5]), float(self.lindep),
                                 ctypes.byref(h)), 'mjk_df_build')
        self.handle = h
        return self

    def __del__(self):
        try:
            if self.handle is not None:
                self.engine.lib.mjk_df_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def get_naoaux(self):
        self.build()
        return int(self.engine.lib.mjk_df_naux(self.handle))

    def cderi(self):
        """B as (naux, nao, nao) (the device holds it as (nao, nao, naux))."""
        self.build()
        nao = self.mol.nao_nr()
        out = numpy.empty((nao, nao, self.get_naoaux()))
        self.engine.check(self.engine.lib.mjk_df_fetch(self.engine.handle, self.handle, 1, _p(out)), 'mjk_df_fetch')
        return numpy.ascontiguousarray(out.transpose(2, 0, 1))

    @property
    def _cderi(self):
        """The reference's layout: (naux, nao(nao+1)/2), lower triangle row by row."""
        b = self.cderi()
        i, j = numpy.tril_indices(b.shape[-1])
        return numpy.ascontiguousarray(b[:, i, j])

    def get_jk(self, dm, hermi=1, with_j=True, with_k=True, direct_scf_tol=None, omega=None):
        from pyscf_amd.df import df_jk
        if omega:
            raise NotImplementedError('range-separated density fitting')
        return df_jk.get_jk(self, dm, hermi, with_j, with_k)
