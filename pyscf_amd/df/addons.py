"""Auxiliary basis selection (pyscf/df/addons.py:41-70, 166-272).

The JK fitting set of an orbital basis follows the reference's table for the
sets shipped here: cc-pVDZ and 6-31G -> cc-pVDZ-JKFIT, def2-SVP and
def2-TZVP -> def2-universal-JKFIT.  Even-tempered generation for other sets
is not built (an explicit error names the basis)."""

from pyscf_amd import gto
from pyscf_amd.gto import basis as _basis

DEFAULT_AUXBASIS = {
    'ccpvdz': 'cc-pvdz-jkfit',
    '631g': 'cc-pvdz-jkfit',
    'def2svp': 'def2-universal-jkfit',
    'def2tzvp': 'def2-universal-jkfit',
}


def default_auxbasis(mol):
    if not isinstance(mol.basis, str):
        raise NotImplementedError('default auxiliary basis for a per-element basis dict')
    key = _basis._format_name(mol.basis)
    if key not in DEFAULT_AUXBASIS:
        raise NotImplementedError('no JK fitting basis shipped for %s' % mol.basis)
    return DEFAULT_AUXBASIS[key]


def make_auxmol(mol, auxbasis=None):
    """Molecule of the auxiliary basis on the same atoms (same units, charge
    and spin; only its basis functions are used)."""
    if auxbasis is None:
        auxbasis = default_auxbasis(mol)
    return gto.M(atom=mol.atom, basis=auxbasis, unit=mol.unit, charge=mol.charge, spin=mol.spin)
