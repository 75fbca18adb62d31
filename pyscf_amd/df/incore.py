"""cholesky_eri (pyscf/df/incore.py:129-214): the fitted three-centre tensor
in the reference's packed layout (naux, nao(nao+1)/2), built on the device."""

from pyscf_amd.df.df import DF


def cholesky_eri(mol, auxbasis=None, auxmol=None):
    if auxmol is not None:
        raise NotImplementedError('an explicit auxmol; pass auxbasis')
    return DF(mol, auxbasis).build()._cderi
