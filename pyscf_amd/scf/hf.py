"""Mean-field front end of the J/K engine (pyscf/scf/hf.py counterparts).

``get_jk`` keeps the reference signature and semantics
(pyscf/scf/hf.py:948-1022): any stack shape, complex densities split into
real and imaginary parts with hermi forced to 0, new float64 outputs,
``None`` for a matrix that was not requested.  ``get_veff`` is
pyscf/scf/hf.py:1025-1075.  The RHF driver lives in ``pyscf_amd.scf.rhf``.
"""

import numpy

from pyscf_amd.scf import _vhf


def get_jk(mol, dm, hermi=1, vhfopt=None, with_j=True, with_k=True, omega=None, fp32_tol=0.0):
    """Coulomb (J) and exchange (K) matrices of one or more densities.

    hermi: 0 general dm, 1 symmetric (Hermitian) dm, 2 anti-symmetric.
    vhfopt: a ``_vhf.VHFOpt`` enables Schwarz x density screening at its
    ``direct_scf_tol``; ``None`` evaluates every quartet.
    omega: None keeps the molecule's operator; 0 the Coulomb 1/r, > 0 the
    long-range erf(omega r)/r, < 0 the short-range erfc(|omega| r)/r
    (``mol.with_range_coulomb``, pyscf/scf/hf.py:1006).  A ``vhfopt`` must
    have been built for the same omega.
    fp32_tol: > 0 evaluates small quartets in FP32 (mixed precision).
    """
    dm = numpy.asarray(dm, order='C')
    dm_shape = dm.shape
    dm_dtype = dm.dtype
    nao = dm_shape[-1]
    if dm_dtype == numpy.complex128:
        dm = numpy.vstack((dm.real, dm.imag)).reshape(-1, nao, nao)
        hermi = 0
    with mol.with_range_coulomb(omega):
        vj, vk = _vhf.direct(dm, mol._atm, mol._bas, mol._env, vhfopt, hermi, getattr(mol, 'cart', False),
                             with_j, with_k, fp32_tol=fp32_tol)
    if dm_dtype == numpy.complex128:
        if with_j:
            vj = vj.reshape((2,) + dm_shape)
            vj = vj[0] + vj[1] * 1j
        if with_k:
            vk = vk.reshape((2,) + dm_shape)
            vk = vk[0] + vk[1] * 1j
    else:
        if with_j:
            vj = vj.reshape(dm_shape)
        if with_k:
            vk = vk.reshape(dm_shape)
    return vj, vk


def get_veff(mol, dm, dm_last=None, vhf_last=None, hermi=1, vhfopt=None):
    """Hartree-Fock potential J - K/2 (incremental when dm_last is given)."""
    if dm_last is None:
        vj, vk = get_jk(mol, numpy.asarray(dm), hermi, vhfopt)
        return vj - vk * .5
    ddm = numpy.asarray(dm) - numpy.asarray(dm_last)
    vj, vk = get_jk(mol, ddm, hermi, vhfopt)
    return vj - vk * .5 + numpy.asarray(vhf_last)
