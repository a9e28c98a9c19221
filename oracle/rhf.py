"""ORACLE (test infrastructure): psi4's RHF SCF loop restated in numpy.

Follows psi4/driver/procrouting/scf_proc/scf_iterator.py:264-609 (iterate),
libscf_solver/rhf.cc:187-367 (form_G, form_F, compute_E), hf.cc:1415-1425
(orbital gradient X^T(FDS-SDF)X), subclass_methods.py:24-45 and diis.py:251-273
(Pulay DIIS on that gradient), scf_iterator.py:929-935 (convergence test).
The guess is the core Hamiltonian; the converged energy does not depend on it.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import eigh


def symmetric_orthogonalizer(S):
    w, U = np.linalg.eigh(S)
    return (U * w ** -0.5) @ U.T


class DIIS:
    def __init__(self, max_vecs=10):
        self.max_vecs = max_vecs
        self.F, self.err = [], []

    def add(self, F, err):
        self.F.append(F.copy())
        self.err.append(err.copy())
        if len(self.F) > self.max_vecs:   # RemovalPolicy.LargestError (diis.py)
            worst = int(np.argmax([np.linalg.norm(e) for e in self.err]))
            del self.F[worst]
            del self.err[worst]

    def extrapolate(self):
        n = len(self.F)
        B = -np.ones((n + 1, n + 1))
        B[n, n] = 0.0
        for i in range(n):
            for j in range(n):
                B[i, j] = np.vdot(self.err[i], self.err[j])
        rhs = np.zeros(n + 1)
        rhs[n] = -1.0
        c = np.linalg.lstsq(B, rhs, rcond=None)[0]
        return sum(c[i] * self.F[i] for i in range(n))


def rhf(H, S, nocc, enuc, jk, e_conv=1e-10, d_conv=1e-8, maxiter=100, diis_max=10, alpha=1.0):
    """jk(C_occ) -> (J, K).  Returns dict(E, C, eps, D, J, K, F, iterations, history)."""
    X = symmetric_orthogonalizer(S)
    eps, C = eigh(H, S)
    diis = DIIS(diis_max)
    E_old = 0.0
    history = []
    for it in range(1, maxiter + 1):
        Cocc = C[:, :nocc]
        D = Cocc @ Cocc.T
        J, K = jk(Cocc)
        F = H + 2.0 * J - alpha * K                                   # rhf.cc:187-243 + form_F
        E = enuc + np.vdot(D, H) * 2.0 + np.vdot(D, 2.0 * J - alpha * K)  # rhf.cc:308-367
        grad = X.T @ (F @ D @ S - S @ D @ F) @ X                      # hf.cc:1415-1425
        d_rms = np.sqrt(np.mean(grad ** 2))
        history.append((E, d_rms))
        if abs(E - E_old) < e_conv and d_rms < d_conv:                # scf_iterator.py:929-935
            return dict(E=E, C=C, eps=eps, D=D, J=J, K=K, F=F, iterations=it, history=history)
        diis.add(F, grad)
        Fx = diis.extrapolate() if it >= 1 else F
        eps, C = eigh(Fx, S)
        E_old = E
    raise RuntimeError(f"SCF did not converge in {maxiter} iterations")
