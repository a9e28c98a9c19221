"""UHF driver on the JK object: two density pairs per JK call.

Restates psi4's UHF (libscf_solver/uhf.cc: form_G J = J_a + J_b, G_s = J - K_s;
form_F; compute_E = E_nuc + 1/2 tr((D_a + D_b) H) + 1/2 sum_s tr(D_s F_s); orbital
gradient X^T (F_s D_s S - S D_s F_s) X of both spins, concatenated for the Pulay
DIIS of procrouting/diis.py) on the in-core DF JK: each iteration is one
``compute()`` with C_left = [C_a,occ, C_b,occ] (lr_symmetric), i.e. two density
pairs through the engine at once.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import eigh

from . import ints
from .core import JK, PREC_FP64
from .scf import SCFConvergenceError, _DIIS


class UHFResult:
    def __init__(self, **kw):
        self.__dict__.update(kw)


class UHF:
    def __init__(self, mol, basis="cc-pvdz", df_basis_scf=None, jk=None, e_convergence=1e-10, d_convergence=1e-8,
                 maxiter=100, diis_max_vecs=10, guess=None):
        from .basisset import BasisSet
        self.mol = mol
        self.basis = BasisSet.build(mol, basis)
        self.aux = BasisSet.build(mol, df_basis_scf or (basis + "-jkfit"))
        nel = mol.nelectron()
        nunp = mol.multiplicity - 1
        if (nel - nunp) % 2 or nunp < 0 or nunp > nel:
            raise ValueError("charge and multiplicity are inconsistent")
        self.nbeta = (nel - nunp) // 2
        self.nalpha = self.nbeta + nunp
        self.e_conv, self.d_conv, self.maxiter = e_convergence, d_convergence, maxiter
        self.diis_max_vecs = diis_max_vecs
        self.guess = guess  # None: core Hamiltonian; or starting orbitals (nbf x nmo) for both spins
        self._jk = jk
        self.Ca = self.Cb = self.Fa = self.Fb = self.eps_a = self.eps_b = None

    def jk(self):
        if self._jk is None:
            self._jk = JK.build(self.basis, self.aux)
            self._jk.set_print(0)
            self._jk.initialize()
        return self._jk

    def _build(self, Ca_occ, Cb_occ):
        jk = self.jk()
        jk.set_precision(PREC_FP64)
        jk.C_clear()
        jk.C_left_add(np.ascontiguousarray(Ca_occ))
        jk.C_left_add(np.ascontiguousarray(Cb_occ))
        jk.compute()
        J, K = jk.J(), jk.K()
        return np.array(J[0]), np.array(J[1]), np.array(K[0]), np.array(K[1])

    def compute_energy(self):
        mol, bs = self.mol, self.basis
        S = ints.overlap(bs)
        H = ints.kinetic(bs) + ints.potential(bs, mol)
        enuc = mol.nuclear_repulsion_energy()
        w, U = np.linalg.eigh(S)
        X = (U * w ** -0.5) @ U.T
        self.jk()
        if self.guess is None:
            eps_a, Ca = eigh(H, S)
        else:
            Ca = np.array(self.guess, dtype=np.float64)
            eps_a = np.zeros(Ca.shape[1])
        eps_b, Cb = eps_a.copy(), Ca.copy()
        diis = _DIIS(self.diis_max_vecs)
        E_old = 0.0
        na, nb = self.nalpha, self.nbeta
        for it in range(1, self.maxiter + 1):
            Ca_o, Cb_o = Ca[:, :na], Cb[:, :nb]
            Da, Db = Ca_o @ Ca_o.T, Cb_o @ Cb_o.T
            Ja, Jb, Ka, Kb = self._build(Ca_o, Cb_o)
            Jt = Ja + Jb
            Fa = H + Jt - Ka
            Fb = H + Jt - Kb
            E = enuc + 0.5 * (np.vdot(Da + Db, H) + np.vdot(Da, Fa) + np.vdot(Db, Fb))
            ga = X.T @ (Fa @ Da @ S - S @ Da @ Fa) @ X
            gb = X.T @ (Fb @ Db @ S - S @ Db @ Fb) @ X
            err = np.concatenate([ga.ravel(), gb.ravel()])
            d_rms = float(np.sqrt(np.mean(err ** 2)))
            if abs(E - E_old) < self.e_conv and d_rms < self.d_conv:
                self.Ca, self.Cb, self.Fa, self.Fb, self.eps_a, self.eps_b = Ca, Cb, Fa, Fb, eps_a, eps_b
                return UHFResult(energy=E, Ca=Ca, Cb=Cb, Da=Da, Db=Db, iterations=it)
            diis.add(np.stack([Fa, Fb]), err)
            Fx = diis.extrapolate()
            eps_a, Ca = eigh(Fx[0], S)
            eps_b, Cb = eigh(Fx[1], S)
            E_old = E
        raise SCFConvergenceError(f"UHF did not converge in {self.maxiter} iterations")
