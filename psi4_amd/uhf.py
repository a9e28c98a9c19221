"""UHF driver on the JK object: two density pairs per JK call.

Restates psi4's UHF (libscf_solver/uhf.cc: form_G J = J_a + J_b, G_s = J - K_s;
form_F; compute_E = E_nuc + 1/2 tr((D_a + D_b) H) + 1/2 sum_s tr(D_s F_s); orbital
gradient X^T (F_s D_s S - S D_s F_s) X of both spins, concatenated for the Pulay
DIIS of procrouting/diis.py) on the in-core DF JK: each iteration is one
``compute()`` with C_left = [C_a,occ, C_b,occ] (lr_symmetric), i.e. two density
pairs through the engine at once.  As in the RHF driver, the SCF algebra (orthogonalised
eigensolves, DIIS, orbital gradient) runs in torch FP64 on the engine's device.

Incremental mixed-precision builds (``mixed_precision``) follow the RHF driver's policy per
spin: after a full FP64 build, each iteration adds J(dD_a) + J(dD_b) and K_s(dD_s) of the
density changes, computed as two cross-form pairs (C_left = [A_s, V_s], C_right = [V_s, A_s]
after a Procrustes alignment of each spin's old orbitals) in one engine call; a full FP64
build every ``incfock_full_every`` iterations, whenever the running sum of the increments'
error bounds c (||dD_a||_F + ||dD_b||_F) would pass ``build_tol``, and for every iteration
once the orbital gradient is below ``mixed_until``.
"""
from __future__ import annotations

import numpy as np
import torch

from . import ints
from .core import JK, PREC_FP64
from .scf import MIXED_ERROR_PER_DD, SCFConvergenceError, _DIIS, _PREC, orthogonalizer


class UHFResult:
    def __init__(self, **kw):
        self.__dict__.update(kw)


class UHF:
    def __init__(self, mol, basis="cc-pvdz", df_basis_scf=None, jk=None, e_convergence=1e-10, d_convergence=1e-8,
                 maxiter=100, diis_max_vecs=10, guess=None, s_tolerance=1e-7, mixed_precision=None,
                 mixed_until=1e-7, incfock_full_every=8, build_tol=1e-6, jk_callback=None):
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
        self.s_tolerance = s_tolerance
        if mixed_precision not in _PREC:
            raise ValueError(f"unknown precision {mixed_precision}")
        self.mixed = None if mixed_precision == "fp64" else mixed_precision
        self.mixed_until = mixed_until
        self.full_every = max(1, int(incfock_full_every))
        self.build_tol = float(build_tol)
        self.history = []
        # called as cb(iteration, Ca_occ, Cb_occ, (J_a, J_b, K_a, K_b), kind) after every build
        self.jk_callback = jk_callback
        self._jk = jk
        self.Ca = self.Cb = self.Fa = self.Fb = self.eps_a = self.eps_b = None

    def jk(self):
        if self._jk is None:
            self._jk = JK.build(self.basis, self.aux)
            self._jk.set_print(0)
            self._jk.initialize()
        return self._jk

    def _device(self):
        """As RHF: the SCF algebra runs on the GPU next to the device engine, else on the host."""
        jk = self.jk()
        if hasattr(jk, "compute_device") and torch.cuda.is_available():
            return torch.device("cuda", torch.cuda.current_device())
        return torch.device("cpu")

    def _delta_build(self, pairs):
        """J and K of the density changes of both spins, one engine call with two cross-form
        pairs: pairs = [(C_new, C_old, S C_new)] per spin; returns dJ_a, dJ_b, dK_a, dK_b."""
        jk = self.jk()
        lefts, rights = [], []
        for Cn, Co, SCn in pairs:
            if Cn.shape[1] == 0:
                lefts.append(Cn.contiguous())
                rights.append(Cn.contiguous())
                continue
            u, _, vt = np.linalg.svd((Co.T @ SCn).cpu().numpy())   # orthogonal Procrustes rotation
            Col = Co @ torch.from_numpy(u @ vt).to(Co.device)       # same density, aligned with Cn
            V = Cn - Col
            A = 0.5 * (Cn + Col)
            lefts.append(torch.cat([A, V], 1).contiguous())
            rights.append(torch.cat([V, A], 1).contiguous())
        jk.set_precision(_PREC[self.mixed])
        jk.set_pair_form("CROSS")
        try:
            n = lefts[0].shape[0]
            out = [torch.empty((n, n), dtype=torch.float64, device=lefts[0].device) for _ in range(4)]
            torch.cuda.synchronize()           # C ready before the engine's stream reads it
            jk.compute_device([L.data_ptr() for L in lefts], [R.data_ptr() for R in rights],
                              [L.shape[1] for L in lefts], [], [out[0].data_ptr(), out[1].data_ptr()],
                              [out[2].data_ptr(), out[3].data_ptr()], False)
            torch.cuda.synchronize()
            return out
        finally:
            jk.set_precision(PREC_FP64)
            jk.set_pair_form("DETECT")

    def _build(self, Ca_occ, Cb_occ):
        """J_a, J_b, K_a, K_b of the two occupied sets, one compute() with two pairs."""
        jk = self.jk()
        jk.set_precision(PREC_FP64)
        if Ca_occ.is_cuda:
            n = Ca_occ.shape[0]
            A, B = Ca_occ.contiguous(), Cb_occ.contiguous()
            out = [torch.empty((n, n), dtype=torch.float64, device=A.device) for _ in range(4)]
            torch.cuda.synchronize()           # C ready before the engine's stream reads it
            jk.compute_device([A.data_ptr(), B.data_ptr()], [], [A.shape[1], B.shape[1]], [],
                              [out[0].data_ptr(), out[1].data_ptr()], [out[2].data_ptr(), out[3].data_ptr()], False)
            torch.cuda.synchronize()
            return out
        jk.C_clear()
        jk.C_left_add(np.ascontiguousarray(Ca_occ.numpy()))
        jk.C_left_add(np.ascontiguousarray(Cb_occ.numpy()))
        jk.compute()
        J, K = jk.J(), jk.K()
        return [torch.from_numpy(np.array(x)) for x in (J[0], J[1], K[0], K[1])]

    def compute_energy(self):
        mol, bs = self.mol, self.basis
        dev = self._device()
        S = torch.from_numpy(ints.overlap(bs)).to(dev)
        H = torch.from_numpy(ints.kinetic(bs) + ints.potential(bs, mol)).to(dev)
        enuc = mol.nuclear_repulsion_energy()
        X = orthogonalizer(S, self.s_tolerance)
        if X.shape[1] < max(self.nalpha, self.nbeta):
            raise RuntimeError("Not enough molecular orbitals to satisfy requested occupancies")

        def diag(F):
            e, Cp = torch.linalg.eigh(X.T @ F @ X)
            return e, X @ Cp

        if self.guess is None:
            eps_a, Ca = diag(H)
        else:
            Ca = torch.as_tensor(np.array(self.guess, dtype=np.float64)).to(dev)
            eps_a = torch.zeros(Ca.shape[1], dtype=torch.float64, device=dev)
        eps_b, Cb = eps_a.clone(), Ca.clone()
        diis = _DIIS(self.diis_max_vecs)
        E_old = 0.0
        na, nb = self.nalpha, self.nbeta
        mixed_on = self.mixed is not None and dev.type == "cuda"
        prev = None                   # (Ca_o, Cb_o, Da, Db) of the last build
        Ja = Jb = Ka = Kb = None
        since_full, err_bound = 0, 0.0
        self.history = []
        for it in range(1, self.maxiter + 1):
            Ca_o, Cb_o = Ca[:, :na], Cb[:, :nb]
            Da, Db = Ca_o @ Ca_o.T, Cb_o @ Cb_o.T
            bound = None
            if prev is not None and mixed_on and since_full + 1 < self.full_every:
                bound = MIXED_ERROR_PER_DD[self.mixed] * (float(torch.linalg.norm(Da - prev[2])) +
                                                          float(torch.linalg.norm(Db - prev[3])))
                if err_bound + bound > self.build_tol:
                    bound = None          # this increment could exceed the per-build bound
            if bound is None:
                Ja, Jb, Ka, Kb = self._build(Ca_o, Cb_o)
                kind, since_full, err_bound = "fp64", 0, 0.0
            else:
                dJa, dJb, dKa, dKb = self._delta_build([(Ca_o, prev[0], S @ Ca_o), (Cb_o, prev[1], S @ Cb_o)])
                Ja, Jb, Ka, Kb = Ja + dJa, Jb + dJb, Ka + dKa, Kb + dKb
                kind = self.mixed
                since_full += 1
                err_bound += bound
            prev = (Ca_o, Cb_o, Da, Db)
            if self.jk_callback is not None:
                self.jk_callback(it, Ca_o.cpu().numpy(), Cb_o.cpu().numpy(),
                                 tuple(x.cpu().numpy() for x in (Ja, Jb, Ka, Kb)), kind)
            Jt = Ja + Jb
            Fa = H + Jt - Ka
            Fb = H + Jt - Kb
            E = enuc + 0.5 * float(torch.sum((Da + Db) * H) + torch.sum(Da * Fa) + torch.sum(Db * Fb))
            FDSa, FDSb = Fa @ Da @ S, Fb @ Db @ S
            ga = X.T @ (FDSa - FDSa.T) @ X
            gb = X.T @ (FDSb - FDSb.T) @ X
            err = torch.cat([ga.reshape(-1), gb.reshape(-1)])
            d_rms = float(torch.sqrt(torch.mean(err * err)))
            self.history.append(dict(iteration=it, energy=E, d_rms=d_rms, build=kind, err_bound=err_bound))
            # convergence is only accepted on an FP64 build
            if abs(E - E_old) < self.e_conv and d_rms < self.d_conv and kind == "fp64":
                h = [x.cpu().numpy() for x in (Ca, Cb, Fa, Fb, eps_a, eps_b, Da, Db)]
                self.Ca, self.Cb, self.Fa, self.Fb, self.eps_a, self.eps_b = h[:6]
                return UHFResult(energy=E, Ca=h[0], Cb=h[1], Da=h[6], Db=h[7], iterations=it)
            if mixed_on and d_rms < self.mixed_until:
                mixed_on = False          # FP64 finishing iterations from here on
            diis.add(torch.stack([Fa, Fb]), err)
            Fx = diis.extrapolate()
            eps_a, Ca = diag(Fx[0])
            eps_b, Cb = diag(Fx[1])
            E_old = E
        raise SCFConvergenceError(f"UHF did not converge in {self.maxiter} iterations")
