"""MI355X-native direct Coulomb/exchange (J/K) engine with PySCF's
``get_jk(mol, dm, hermi) -> (vj, vk)`` interface."""
__version__ = '0.1.0'
