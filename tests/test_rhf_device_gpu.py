"""Device-resident RHF (mjk_rhf_*) against the host-driven RHF of the same
engine: the two loops take the same steps (core guess, CDIIS from cycle 1,
incremental Fock builds) with the dense algebra on rocBLAS/rocSOLVER in
place of numpy/LAPACK, so they converge to the same state:

* energies within 1e-9 Eh at conv_tol 1e-11 (both converged), densities
  within 1e-6, occupied orbital energies within 1e-6 Eh, cycle counts
  within two;
* with a level shift and with a caller's initial density;
* the host RHF is pinned against the oracle's J/K and the reference's
  energies in the SCF tests; this test ties the device loop to it.
"""
import numpy
import pytest

from pyscf_amd import gto, workloads
from pyscf_amd.scf import rhf

pytestmark = pytest.mark.gpu

MOLS = {
    'h2o': lambda: gto.M(atom=workloads.h2o(), basis='cc-pvdz'),
    'water4': lambda: gto.M(atom=workloads.water_cluster(4, seed=3), basis='def2-svp'),
    'butane': lambda: gto.M(atom=workloads.alkane(4), basis='cc-pvdz'),
}


def run(cls, mol, dm0=None, **attrs):
    mf = cls(mol, direct_scf_tol=1e-13)
    mf.conv_tol = 1e-11
    for k, v in attrs.items():
        setattr(mf, k, v)
    mf.kernel(dm0)
    return mf


@pytest.mark.parametrize('name', sorted(MOLS))
def test_device_rhf_matches_host(name):
    mol = MOLS[name]()
    host = run(rhf.RHF, mol)
    dev = run(rhf.DeviceRHF, mol)
    nocc = mol.nelectron // 2
    print('%s: host E=%.12f (%d cycles) device E=%.12f (%d cycles) |g|=%.2e'
          % (name, host.e_tot, host.cycles, dev.e_tot, dev.cycles, dev.gnorm), flush=True)
    assert host.converged and dev.converged
    assert abs(host.e_tot - dev.e_tot) <= 1e-9
    assert abs(host.dm - dev.dm).max() <= 1e-6
    assert abs(numpy.sort(host.mo_energy)[:nocc] - dev.mo_energy[:nocc]).max() <= 1e-6
    assert abs(host.cycles - dev.cycles) <= 2
    # the fetched orbitals reproduce the fetched density and are S-orthonormal
    c = dev.mo_coeff[:, :nocc]
    assert abs(2 * c @ c.T - dev.dm).max() <= 1e-10
    s = rhf.int1e(mol, 'ovlp')
    assert abs(dev.mo_coeff.T @ s @ dev.mo_coeff - numpy.eye(mol.nao_nr())).max() <= 1e-8


def test_device_rhf_level_shift_and_guess():
    mol = MOLS['water4']()
    ref = run(rhf.RHF, mol)
    shifted = run(rhf.DeviceRHF, mol, level_shift=0.3)
    assert shifted.converged and abs(shifted.e_tot - ref.e_tot) <= 1e-9
    # restart from the converged density: converges in a cycle or two
    mf = run(rhf.DeviceRHF, mol, dm0=ref.dm)
    assert mf.converged and mf.cycles <= 3 and abs(mf.e_tot - ref.e_tot) <= 1e-9


def test_device_rhf_refusals():
    mol = gto.M(atom='H 0 0 0; H 0 0 0.74; H 0 0 1.6', basis='cc-pvdz', spin=1)
    with pytest.raises(ValueError):
        rhf.DeviceRHF(mol).kernel()
