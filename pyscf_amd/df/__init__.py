"""Density fitting (pyscf/df): the fitted three-centre tensor built and kept
on the device, and the J/K builds from it.

* ``DF``            counterpart of pyscf.df.DF (pyscf/df/df.py): ``auxbasis``,
                    ``auxmol``, ``build()``, ``get_jk()``, ``get_naoaux()``;
                    ``_cderi`` gives the reference's packed (naux, nao(nao+1)/2)
                    array on demand.
* ``make_auxmol``   the auxiliary molecule (pyscf/df/addons.py:226-272),
                    default fitting sets of pyscf/df/addons.py:41-70 for the
                    orbital sets shipped here.
* ``density_fit``   the density-fitted SCF object (pyscf/df/df_jk.py:32-125).
"""

from pyscf_amd.df.addons import DEFAULT_AUXBASIS, make_auxmol
from pyscf_amd.df.df import DF
from pyscf_amd.df.df_jk import density_fit, get_jk
from pyscf_amd.df import incore
