"""Molecule/basis model producing PySCF's ``_atm/_bas/_env`` arrays."""
from pyscf_amd.gto.mole import Mole, M, BOHR, gto_norm
