"""Mean-field layer: get_jk / get_veff front end and the RHF driver."""
from pyscf_amd.scf import hf, _vhf
from pyscf_amd.scf.hf import get_jk, get_veff
