"""MI355X-native density-fitted J/K (MemDFJK) engine with psi4's JK object API.

Submodules
  molecule   Molecule, Angstrom/bohr, seeded water-cluster generator
  basisset   psi4 .gbs basis sets flattened for the C ABI
  ints       host integral engine bindings (S, T, V, (P|Q), (P|mn), Schwarz)
  core       JK.build / JK objects (pybind11 over include/psi4amd_dfjk.h)
  scf        RHF driver on the JK object (psi4 scf_iterator restated)
"""
__version__ = "0.1.0"
