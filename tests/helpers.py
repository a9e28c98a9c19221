"""Shared builders for tests (molecules, oracle JK, GPU JK)."""
import numpy as np

from psi4_amd.basisset import BasisSet
from psi4_amd.molecule import water_cluster


def water_system(grid=(2, 2, 1), seed=0, basis="cc-pvdz", aux="cc-pvdz-jkfit"):
    mol = water_cluster(grid, seed=seed)
    return mol, BasisSet.build(mol, basis), BasisSet.build(mol, aux)


def random_orbitals(nbf, ncol, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nbf, ncol)) / np.sqrt(nbf)


def gpu_jk(prim, aux, **kw):
    from psi4_amd.core import JK
    jk = JK.build(prim, aux, **kw)
    jk.set_print(0)
    jk.initialize()
    return jk


def maxabs(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))
