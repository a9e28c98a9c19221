"""RHF driver on the JK object: psi4's SCF loop restated as the caller of the DF-JK path.

Restates psi4/driver/procrouting/scf_proc/scf_iterator.py:264-609 (iterate),
libscf_solver/rhf.cc:187-367 (form_G: G = 2J - K, form_F, compute_E),
hf.cc:1415-1425 (orbital gradient X^T (FDS - SDF) X), the Pulay DIIS of
procrouting/diis.py:251-273, the convergence test of scf_iterator.py:929-935,
and the JK injection of proc.py:2043-2045 (an injected JK is used as is and
not re-initialized, scf_iterator.py:153-159).

Incremental mixed-precision Fock builds (``mixed_precision``): after a full
FP64 build, each iteration adds J(dD), K(dD) of the density change to FP64 J
and K, with K(dD) in FP32 or 3-term BF16 on the device.  dD is passed in the
cross form dD = A V^T + V A^T (C_left=[A, V], C_right=[V, A], A = (C_new +
C_old)/2, V = C_new - C_old, C_old aligned to C_new by an orthogonal Procrustes
rotation), so every low-precision product carries the small factor V and K(dD)
stays symmetric.  A full FP64 rebuild runs
every ``incfock_full_every`` iterations (psi4's INCFOCK_FULL_FOCK_EVERY
pattern, libfock/DirectJK.cc:333-349), and once the orbital gradient falls
below ``mixed_until`` every remaining iteration is a full FP64 build.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
from scipy.linalg import eigh

from . import ints
from .core import JK, PREC_BF16X3, PREC_FP32, PREC_FP64

_PREC = {None: PREC_FP64, "fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3}


class SCFConvergenceError(RuntimeError):
    """Raised when MAXITER is reached (scf_iterator.py:608-609)."""


@dataclass
class SCFResult:
    energy: float
    C: np.ndarray
    eps: np.ndarray
    D: np.ndarray
    J: np.ndarray
    K: np.ndarray
    F: np.ndarray
    iterations: int
    history: list = field(default_factory=list)
    timings: dict = field(default_factory=dict)


class _DIIS:
    def __init__(self, max_vecs):
        self.max_vecs = max_vecs
        self.F, self.err = [], []

    def add(self, F, err):
        self.F.append(F.copy())
        self.err.append(err.copy())
        if len(self.F) > self.max_vecs:
            worst = int(np.argmax([np.linalg.norm(e) for e in self.err]))
            del self.F[worst]
            del self.err[worst]

    def extrapolate(self):
        n = len(self.F)
        B = -np.ones((n + 1, n + 1))
        B[n, n] = 0.0
        for i in range(n):
            for j in range(i, n):
                B[i, j] = B[j, i] = np.vdot(self.err[i], self.err[j])
        rhs = np.zeros(n + 1)
        rhs[n] = -1.0
        c = np.linalg.lstsq(B, rhs, rcond=None)[0]
        return sum(c[i] * self.F[i] for i in range(n))


class RHF:
    """Closed-shell SCF with a DF JK object (MEM_DF path)."""

    def __init__(self, mol, basis="cc-pvdz", df_basis_scf=None, jk=None, e_convergence=1e-10, d_convergence=1e-8,
                 maxiter=100, diis_max_vecs=10, mixed_precision=None, mixed_until=1e-5, incfock_full_every=5,
                 fail_on_maxiter=True, jk_callback=None, perturb_dipole=None):
        from .basisset import BasisSet
        self.mol = mol
        self.basis = BasisSet.build(mol, basis)
        self.aux = BasisSet.build(mol, df_basis_scf or (basis + "-jkfit"))
        self.nocc = mol.nelectron() // 2
        if mol.nelectron() % 2 or mol.multiplicity != 1:
            raise ValueError("RHF needs a closed-shell singlet")
        self.e_conv, self.d_conv, self.maxiter = e_convergence, d_convergence, maxiter
        self.diis_max_vecs = diis_max_vecs
        self.mixed = mixed_precision
        if mixed_precision not in _PREC:
            raise ValueError(f"unknown precision {mixed_precision}")
        self.mixed_until = mixed_until
        self.full_every = max(1, int(incfock_full_every))
        self.fail_on_maxiter = fail_on_maxiter
        self.jk_callback = jk_callback  # called as cb(iteration, C_occ, J, K, kind) after every build
        self._jk = jk
        self._injected = jk is not None
        self.Ca = self.Fa = self.eps = None
        # PERTURB_H / PERTURB_WITH DIPOLE (libmints/mintshelper.cc:1433-1480): V += sum_d lambda_d mu_d
        # with psi4's electronic dipole integrals mu_d = -<r_d>
        self.perturb_dipole = None if perturb_dipole is None else np.asarray(perturb_dipole, np.float64)
        if jk is not None and jk.info()["nbf"] != self.basis.nbf():        # hf.cc:400-409
            raise ValueError("JK object basis does not match the wavefunction basis")

    # psi4 _build_jk / initialize_jk (scf_iterator.py:103-129)
    def jk(self):
        if self._jk is None:
            self._jk = JK.build(self.basis, self.aux)
            self._jk.set_print(0)
            self._jk.initialize()
        return self._jk

    # CPHF products of the converged reference (rhf.cc:372-842), restated in response.py
    def onel_Hx(self, x_vec):
        from . import response
        return response.onel_Hx(self, x_vec)

    def twoel_Hx_full(self, x_vec, combine=True, return_basis="MO", singlet=True):
        from . import response
        return response.twoel_Hx_full(self, x_vec, combine, return_basis, singlet)

    def cphf_Hx(self, x_vec):
        from . import response
        return response.cphf_Hx(self, x_vec)

    def cphf_solve(self, x_vec, conv_tol=1e-5, max_iter=10, print_lvl=0):
        from . import response
        return response.cphf_solve(self, x_vec, conv_tol, max_iter, print_lvl)

    def _full_build(self, Cocc):
        jk = self.jk()
        jk.set_precision(PREC_FP64)
        jk.C_clear()
        jk.C_left_add(Cocc)
        jk.compute()
        return jk.J()[0].copy(), jk.K()[0].copy()

    def _delta_build(self, Cnew, Cold, S):
        """J(dD), K(dD) for dD = Cnew Cnew^T - Cold Cold^T in the cross form
        dD = A V^T + V A^T, A = (Cnew + Cold)/2, V = Cnew - Cold (C_left = [A, V],
        C_right = [V, A]), after aligning Cold to Cnew: every low-precision product
        carries the small factor V, and the engine keeps K(dD) symmetric."""
        u, _, vt = np.linalg.svd(Cold.T @ S @ Cnew)
        Col = Cold @ (u @ vt)            # same density, aligned with Cnew
        V = Cnew - Col
        A = 0.5 * (Cnew + Col)
        jk = self.jk()
        jk.set_precision(_PREC[self.mixed])
        jk.C_clear()
        jk.C_left_add(np.ascontiguousarray(np.hstack([A, V])))
        jk.C_right_add(np.ascontiguousarray(np.hstack([V, A])))
        jk.compute()
        jk.set_precision(PREC_FP64)
        return jk.J()[0].copy(), jk.K()[0].copy()

    def compute_energy(self):
        t0 = time.perf_counter()
        mol, bs = self.mol, self.basis
        S = ints.overlap(bs)
        H = ints.kinetic(bs) + ints.potential(bs, mol)
        if self.perturb_dipole is not None:
            H = H - np.einsum("d,dmn->mn", self.perturb_dipole, ints.dipole(bs))
        enuc = mol.nuclear_repulsion_energy()
        w, U = np.linalg.eigh(S)
        X = (U * w ** -0.5) @ U.T
        self.jk()
        t_setup = time.perf_counter() - t0
        eps, C = eigh(H, S)
        diis = _DIIS(self.diis_max_vecs)
        E_old = 0.0
        history = []
        J = K = None
        C_prev = None
        since_full = 0
        mixed_on = self.mixed is not None
        t_jk = 0.0
        for it in range(1, self.maxiter + 1):
            Cocc = np.ascontiguousarray(C[:, :self.nocc])
            D = Cocc @ Cocc.T
            tj = time.perf_counter()
            if J is None or not mixed_on or since_full + 1 >= self.full_every:
                J, K = self._full_build(Cocc)
                kind = "fp64"
                since_full = 0
            else:
                dJ, dK = self._delta_build(Cocc, C_prev, S)
                J = J + dJ
                K = K + dK
                kind = self.mixed
                since_full += 1
            dt_jk = time.perf_counter() - tj
            t_jk += dt_jk
            if self.jk_callback is not None:
                self.jk_callback(it, Cocc, J, K, kind)
            F = H + 2.0 * J - K
            E = enuc + 2.0 * np.vdot(D, H) + np.vdot(D, 2.0 * J - K)
            grad = X.T @ (F @ D @ S - S @ D @ F) @ X
            d_rms = float(np.sqrt(np.mean(grad ** 2)))
            history.append(dict(iteration=it, energy=E, d_rms=d_rms, build=kind, jk_s=dt_jk))
            converged = abs(E - E_old) < self.e_conv and d_rms < self.d_conv
            if converged and kind == "fp64":
                self.Ca, self.Fa, self.eps = C, F, eps
                return SCFResult(E, C, eps, D, J, K, F, it, history,
                                 dict(setup_s=t_setup, jk_s=t_jk, total_s=time.perf_counter() - t0))
            if mixed_on and d_rms < self.mixed_until:
                mixed_on = False         # FP64 finishing iterations from here on
            diis.add(F, grad)
            eps, C = eigh(diis.extrapolate(), S)
            C_prev = Cocc
            E_old = E
        if self.fail_on_maxiter:
            raise SCFConvergenceError(f"SCF did not converge in {self.maxiter} iterations")
        return SCFResult(E, C, eps, D, J, K, F, self.maxiter, history,
                         dict(setup_s=t_setup, jk_s=t_jk, total_s=time.perf_counter() - t0))


def energy(mol, basis="cc-pvdz", jk=None, **kw):
    """psi4.energy('scf') for an RHF closed shell on the MEM_DF path (driver.py:85, proc.py:2737)."""
    return RHF(mol, basis, jk=jk, **kw).compute_energy().energy
