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


class OracleJK:
    """The oracle MemDFJK behind the JK-object methods the RHF and CPHF drivers call
    (C_clear / C_left_add / C_right_add / compute / J / K / D), for CPU runs of the
    drivers and as the reference side of GPU driver tests."""

    def __init__(self, prim, aux, **kw):
        from oracle.dfhelper import MemDFJKOracle
        self.ora = MemDFJKOracle(prim, aux, **kw)
        self._L, self._R = [], []
        self._D = self._J = self._K = []

    def info(self):
        return {"nbf": self.ora.nbf, "naux": self.ora.naux}

    def set_print(self, _):
        pass

    def set_precision(self, _):
        pass

    def initialize(self):
        pass

    def C_clear(self):
        self._L, self._R = [], []

    def C_left_add(self, C):
        self._L.append(np.asarray(C))

    def C_right_add(self, C):
        self._R.append(np.asarray(C))

    def compute(self):
        self._D, self._J, self._K = self.ora.compute(self._L, self._R or None)

    def D(self):
        return self._D

    def J(self):
        return self._J

    def K(self):
        return self._K


def per_build_checker(jk, worst):
    """jk_callback that checks every mixed incremental build against a plain FP64 build of the
    same orbitals: |E_build(mixed) - E_build(FP64)| with E_build = tr(D(2J - K))."""
    def cb(it, C, J, K, kind):
        if kind == "fp64":
            return
        jk.C_clear()
        jk.C_left_add(C)
        jk.compute()
        D = C @ C.T
        e_mixed = float(np.sum(D * (2.0 * J - K)))
        e_ref = float(np.sum(D * (2.0 * jk.J()[0] - jk.K()[0])))
        worst.append(abs(e_mixed - e_ref))
        print(f"  iteration {it} ({kind}): |dE_build| {worst[-1]:.2e}", flush=True)
    return cb
