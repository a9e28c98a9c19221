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

Per-build error bound: the energy error of a low-precision increment grows with
the size of the density change, |dE| <= c_prec ||dD||_F (c_prec measured on
water-cluster SCF trajectories, MIXED_ERROR_PER_DD), and the increments since the
last FP64 build add up in J and K.  An increment is therefore taken in low
precision only while the running sum of these bounds stays within ``build_tol``
(1e-6 Eh); otherwise the iteration is a full FP64 build, which resets the sum.
Every build's J and K are thus within ``build_tol`` of the FP64 build of the
same orbitals.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import ints
from .core import JK, PREC_BF16X3, PREC_FP32, PREC_FP64

_PREC = {None: PREC_FP64, "fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3}

# Bound on |E_build(mixed) - E_build(FP64)| per unit ||dD||_F of one cross-form increment
# J(dD), K(dD) (Eh).  Measured along the plain FP64 SCF trajectory of the seeded (H2O)_48
# cluster (cc-pVDZ/cc-pVDZ-jkfit, 23 increments, ||dD||_F from 100 down to 1e-5): the ratio
# never exceeded 4.1e-6 for BF16x3 (K dominated; J reads the FP32 copy of B) and 6.1e-7 for
# FP32; the synthetic nbf=1500 increment gives 8e-6 for BF16x3.  The bounds below keep a
# margin above all of them.
MIXED_ERROR_PER_DD = {"fp32": 2e-6, "bf16x3": 1e-5}


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


def _copy(a):
    return a.clone() if torch.is_tensor(a) else a.copy()


def orthogonalizer(S, s_tolerance=1e-7):
    """X with X^T S X = 1: psi4's S_ORTHOGONALIZATION AUTO (libmints/orthog.cc:58-77, 176-235,
    330-360; S_TOLERANCE default 1e-7, read_options.cc:1497).  S is normalised to unit
    diagonal, Sn = U s U^T; if min(s) >= s_tolerance the symmetric X = U s^-1/2 U^T is used,
    else the canonical X = U_k s_k^-1/2 over the eigenvalues above the tolerance (columns in
    increasing eigenvalue order, nmo < nbf); the rows are then scaled back by the
    normalisation.  psi4's third choice, partial Cholesky for a pathologically overcomplete
    basis (reciprocal condition <= machine epsilon), is not provided: that case raises."""
    n = 1.0 / torch.sqrt(torch.diagonal(S))
    Sn = S * n[:, None] * n[None, :]
    s, U = torch.linalg.eigh(Sn)
    rcond = float(s[0] / s[-1])
    if not (rcond > torch.finfo(S.dtype).eps and np.isfinite(rcond)):
        raise RuntimeError("overlap matrix is pathologically overcomplete (partial Cholesky "
                           "orthogonalization is not provided)")
    if float(s[0]) < s_tolerance:
        keep = s > s_tolerance
        X = U[:, keep] * s[keep].rsqrt()
    else:
        X = (U * s.rsqrt()) @ U.T
    return X * n[:, None]


class _DIIS:
    """Pulay DIIS over numpy arrays or torch tensors (host or device).  The error-vector Gram
    matrix is kept between calls: a new vector adds one row of inner products."""

    def __init__(self, max_vecs):
        self.max_vecs = max_vecs
        self.F, self.err = [], []
        self.G = np.zeros((0, 0))

    @staticmethod
    def _dots(vs, e):
        """<v, e> for every stored v (device reductions, one sync for tensors)."""
        if torch.is_tensor(e):
            return torch.stack([torch.sum(v * e) for v in vs]).cpu().numpy()
        return np.array([float(np.vdot(v, e)) for v in vs])

    def add(self, F, err):
        self.F.append(_copy(F))
        self.err.append(_copy(err))
        row = self._dots(self.err, self.err[-1])
        n = len(self.err)
        G = np.empty((n, n))
        G[:n - 1, :n - 1] = self.G
        G[n - 1, :] = row
        G[:, n - 1] = row
        self.G = G
        if n > self.max_vecs:
            worst = int(np.argmax(np.diag(self.G)))
            del self.F[worst]
            del self.err[worst]
            self.G = np.delete(np.delete(self.G, worst, axis=0), worst, axis=1)

    def extrapolate(self):
        n = len(self.F)
        B = -np.ones((n + 1, n + 1))
        B[n, n] = 0.0
        B[:n, :n] = self.G
        rhs = np.zeros(n + 1)
        rhs[n] = -1.0
        c = np.linalg.lstsq(B, rhs, rcond=None)[0]
        return sum(c[i] * self.F[i] for i in range(n))


class RHF:
    """Closed-shell SCF with a DF JK object (MEM_DF path)."""

    def __init__(self, mol, basis="cc-pvdz", df_basis_scf=None, jk=None, e_convergence=1e-10, d_convergence=1e-8,
                 maxiter=100, diis_max_vecs=10, mixed_precision=None, mixed_until=1e-7, incfock_full_every=8,
                 fail_on_maxiter=True, jk_callback=None, perturb_dipole=None, s_tolerance=1e-7, build_tol=1e-6):
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
        self.build_tol = float(build_tol)
        self.full_every = max(1, int(incfock_full_every))
        self.fail_on_maxiter = fail_on_maxiter
        self.jk_callback = jk_callback  # called as cb(iteration, C_occ, J, K, kind) after every build
        self.s_tolerance = s_tolerance
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

    def _device(self):
        """Where the SCF algebra runs: on the GPU next to the engine's J/K when the JK is the
        device engine, else on the host (a JK object that takes and returns numpy arrays)."""
        jk = self.jk()
        if hasattr(jk, "compute_device") and torch.cuda.is_available():
            return torch.device("cuda", torch.cuda.current_device())
        return torch.device("cpu")

    def _jk_call(self, prec, Cl, Cr=None, form=None):
        """J, K of one density pair (C_left, C_right) as tensors on the SCF device.  form
        declares the pair's structure to the engine ("CROSS"), which then skips its test."""
        jk = self.jk()
        jk.set_precision(prec)
        declared = form is not None and hasattr(jk, "set_pair_form")
        if declared:
            jk.set_pair_form(form)
        try:
            if Cl.is_cuda:
                nbf, ncol = Cl.shape
                J = torch.empty((nbf, nbf), dtype=torch.float64, device=Cl.device)
                K = torch.empty_like(J)
                torch.cuda.synchronize()       # C ready before the engine's stream reads it
                jk.compute_device([Cl.data_ptr()], [] if Cr is None else [Cr.data_ptr()], [ncol], [],
                                  [J.data_ptr()], [K.data_ptr()], False)
                torch.cuda.synchronize()       # the engine's stream is done with J, K
                return J, K
            jk.C_clear()
            jk.C_left_add(Cl.numpy())
            if Cr is not None:
                jk.C_right_add(Cr.numpy())
            jk.compute()
            return torch.from_numpy(jk.J()[0].copy()), torch.from_numpy(jk.K()[0].copy())
        finally:
            jk.set_precision(PREC_FP64)
            if declared:
                jk.set_pair_form("DETECT")

    def _full_build(self, Cocc):
        return self._jk_call(PREC_FP64, Cocc)

    def _delta_build(self, Cnew, Cold, SCnew):
        """J(dD), K(dD) for dD = Cnew Cnew^T - Cold Cold^T in the cross form
        dD = A V^T + V A^T, A = (Cnew + Cold)/2, V = Cnew - Cold (C_left = [A, V],
        C_right = [V, A]), after aligning Cold to Cnew: every low-precision product
        carries the small factor V, and the engine keeps K(dD) symmetric.  SCnew = S Cnew."""
        M = (Cold.T @ SCnew).cpu().numpy()
        u, _, vt = np.linalg.svd(M)      # orthogonal Procrustes rotation (nocc x nocc, on the host)
        Col = Cold @ torch.from_numpy(u @ vt).to(Cold.device)   # same density, aligned with Cnew
        V = Cnew - Col
        A = 0.5 * (Cnew + Col)
        return self._jk_call(_PREC[self.mixed], torch.cat([A, V], 1).contiguous(),
                             torch.cat([V, A], 1).contiguous(), form="CROSS")

    def compute_energy(self):
        t0 = time.perf_counter()
        mol, bs = self.mol, self.basis
        S_h = ints.overlap(bs)
        H_h = ints.kinetic(bs) + ints.potential(bs, mol)
        if self.perturb_dipole is not None:
            H_h = H_h - np.einsum("d,dmn->mn", self.perturb_dipole, ints.dipole(bs))
        enuc = mol.nuclear_repulsion_energy()
        dev = self._device()
        S = torch.from_numpy(S_h).to(dev)
        H = torch.from_numpy(H_h).to(dev)
        # orthogonalizer X (symmetric, or canonical when S is near-singular); F C = S C eps is
        # solved as (X^T F X) C' = C' eps, C = X C' (hf.cc form_Shalf / form_C)
        X = orthogonalizer(S, self.s_tolerance)
        if X.shape[1] < self.nocc:
            raise RuntimeError("Not enough molecular orbitals to satisfy requested occupancies")

        def diag(F):
            e, Cp = torch.linalg.eigh(X.T @ F @ X)
            return e, X @ Cp

        self.jk()
        t_setup = time.perf_counter() - t0
        eps, C = diag(H)
        diis = _DIIS(self.diis_max_vecs)
        E_old = 0.0
        history = []
        J = K = None
        C_prev = D_prev = None
        since_full = 0
        err_bound = 0.0          # sum of the increments' error bounds since the last FP64 build
        mixed_on = self.mixed is not None
        t_jk = 0.0
        for it in range(1, self.maxiter + 1):
            Cocc = C[:, :self.nocc].contiguous()
            D = Cocc @ Cocc.T
            tj = time.perf_counter()
            bound = None
            if J is not None and mixed_on and since_full + 1 < self.full_every:
                bound = MIXED_ERROR_PER_DD[self.mixed] * float(torch.linalg.norm(D - D_prev))
                if err_bound + bound > self.build_tol:
                    bound = None         # this increment could exceed the per-build bound
            if bound is None:
                J, K = self._full_build(Cocc)
                kind = "fp64"
                since_full = 0
                err_bound = 0.0
            else:
                dJ, dK = self._delta_build(Cocc, C_prev, S @ Cocc)
                J = J + dJ
                K = K + dK
                kind = self.mixed
                since_full += 1
                err_bound += bound
            dt_jk = time.perf_counter() - tj
            t_jk += dt_jk
            if self.jk_callback is not None:
                self.jk_callback(it, Cocc.cpu().numpy(), J.cpu().numpy(), K.cpu().numpy(), kind)
            F = H + 2.0 * J - K
            E = enuc + float(torch.sum(D * (2.0 * H + 2.0 * J - K)))
            FDS = F @ D @ S
            grad = X.T @ (FDS - FDS.T) @ X
            d_rms = float(torch.sqrt(torch.mean(grad * grad)))
            history.append(dict(iteration=it, energy=E, d_rms=d_rms, build=kind, jk_s=dt_jk, err_bound=err_bound))
            converged = abs(E - E_old) < self.e_conv and d_rms < self.d_conv
            if converged and kind == "fp64":
                return self._result(E, C, eps, D, J, K, F, it, history, t_setup, t_jk, t0)
            if mixed_on and d_rms < self.mixed_until:
                mixed_on = False         # FP64 finishing iterations from here on
            diis.add(F, grad)
            eps, C = diag(diis.extrapolate())
            C_prev, D_prev = Cocc, D
            E_old = E
        if self.fail_on_maxiter:
            raise SCFConvergenceError(f"SCF did not converge in {self.maxiter} iterations")
        return self._result(E, C, eps, D, J, K, F, self.maxiter, history, t_setup, t_jk, t0)

    def _result(self, E, C, eps, D, J, K, F, it, history, t_setup, t_jk, t0):
        h = [x.cpu().numpy() for x in (C, eps, D, J, K, F)]
        self.Ca, self.Fa, self.eps = h[0], h[5], h[1]
        return SCFResult(E, *h, it, history, dict(setup_s=t_setup, jk_s=t_jk, total_s=time.perf_counter() - t0))


def energy(mol, basis="cc-pvdz", jk=None, **kw):
    """psi4.energy('scf') for an RHF closed shell on the MEM_DF path (driver.py:85, proc.py:2737)."""
    return RHF(mol, basis, jk=jk, **kw).compute_energy().energy
