"""RHF driven by the GPU J/K engine reproduces the reference energies."""
import numpy
import pytest

from pyscf_amd import gto, lib
from pyscf_amd.scf import rhf
from tests.conftest import H2O

pytestmark = pytest.mark.gpu


def test_pin7_h2o_631g_energy(pins, h2o_631g):
    mf = rhf.RHF(h2o_631g)
    mf.conv_tol = 1e-11
    e = mf.kernel()
    assert mf.converged
    assert abs(e - pins['pin7_h2o_631g_rhf']['value']) < 1e-9


def test_pin6_h2o_ccpvdz_direct_energy(pins, h2o_vdz):
    mf = rhf.RHF(h2o_vdz)
    mf.conv_tol = 1e-11
    e = mf.kernel()
    assert abs(e - pins['pin6_h2o_ccpvdz_rhf_direct']['value']) < 1e-9


def test_pin8_converged_vj_fingerprint(pins):
    from pyscf_amd.scf import hf
    mol = gto.M(atom=H2O, basis='631g', cart=True)
    mf = rhf.RHF(mol)
    mf.conv_tol = 1e-10
    mf.kernel()
    vj, _ = hf.get_jk(mol, mf.dm)
    assert abs(lib.fp(vj) - pins['pin8_h2o_631g_cart_fp_vj_converged']['value']) < 1e-6
