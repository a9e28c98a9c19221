"""Coupled-perturbed RHF on the DF JK: the CPHF callers of the non-symmetric JK path.

Restates, for a C1 closed-shell reference:
  RHF::onel_Hx              libscf_solver/rhf.cc:372-425
  RHF::twoel_Hx_full        libscf_solver/rhf.cc:427-600   (C_left = C_occ, C_right = -C_vir x^T
                                                             per vector, one JK call, 4J - K - K^T)
  RHF::cphf_Hx              libscf_solver/rhf.cc:602-612
  RHF::cphf_solve           libscf_solver/rhf.cc:613-842   (preconditioned CG, per-vector
                                                             convergence on the relative residual RMS)
  cpscf_linear_response     procrouting/response/scf_response.py:55-181 (dipole polarizabilities)

All products run through the JK object of the converged RHF (``RHF.jk()``):
every trial vector is one (C_left, C_right) pair of a single ``compute()``, so
the multi-pair non-symmetric path of the engine carries the whole response
solve.  Host algebra is FP64 numpy.
"""
from __future__ import annotations

import numpy as np

from . import ints


class CPHFConvergenceError(RuntimeError):
    pass


def _wfn_state(wfn):
    if getattr(wfn, "Ca", None) is None:
        raise RuntimeError("RHF: compute_energy() must run before the response products")
    return wfn.Ca[:, :wfn.nocc], wfn.Ca[:, wfn.nocc:], wfn.Fa


def _check_shapes(wfn, x_vec, who):
    nvir = wfn.Ca.shape[1] - wfn.nocc
    for x in x_vec:
        if x.shape != (wfn.nocc, nvir):
            raise ValueError(f"SCF::{who} incoming rotation matrices must have shape (occ x vir).")


def onel_Hx(wfn, x_vec):
    """F_oo x - x F_vv for each x (occ x vir), rhf.cc:372-425."""
    _check_shapes(wfn, x_vec, "onel_Hx")
    Co, Cv, F = _wfn_state(wfn)
    Foo = Co.T @ F @ Co
    Fvv = Cv.T @ F @ Cv
    return [Foo @ x - x @ Fvv for x in x_vec]


def twoel_Hx_full(wfn, x_vec, combine=True, return_basis="MO", singlet=True):
    """Two-electron part of the RHF orbital Hessian product, rhf.cc:427-600 (HF: alpha = 1)."""
    _check_shapes(wfn, x_vec, "twoel_Hx")
    if return_basis not in ("MO", "SO"):
        raise ValueError("SCF::twoel_Hx: return_basis option not understood.")
    Co, Cv, _ = _wfn_state(wfn)
    jk = wfn.jk()
    jk.set_precision(0)
    jk.C_clear()
    Co = np.ascontiguousarray(Co)
    for x in x_vec:
        jk.C_left_add(Co)
        jk.C_right_add(np.ascontiguousarray(-(Cv @ x.T)))
    jk.compute()
    Js = [np.array(j, copy=True) for j in jk.J()]
    Ks = [np.array(k, copy=True) for k in jk.K()]
    jk.C_clear()
    if not singlet:
        Js = [np.zeros_like(j) for j in Js]
    ret = []
    if combine:
        for J, K in zip(Js, Ks):
            ret.append(4.0 * J - K - K.T)
    else:
        for J, K in zip(Js, Ks):
            ret += [J, K]
    if return_basis == "MO":
        ret = [Co.T @ r @ Cv for r in ret]
    return ret


def cphf_Hx(wfn, x_vec):
    """rhf.cc:602-612."""
    one = onel_Hx(wfn, x_vec)
    two = twoel_Hx_full(wfn, x_vec, True, "MO")
    return [a + b for a, b in zip(one, two)]


def cphf_solve(wfn, x_vec, conv_tol=1e-5, max_iter=10, print_lvl=0, raise_on_fail=False):
    """Preconditioned CG for cphf_Hx(x) = b per right-hand side, rhf.cc:613-842.

    The preconditioner is f_aa - f_ii of the MO Fock matrix; the initial guess is
    b / precon; a vector is converged when sqrt(|r|^2 / |b|^2) < conv_tol.
    Returns the solutions; ``wfn.cphf_converged`` and ``wfn.cphf_nfock_builds``
    record the outcome as psi4 does."""
    _check_shapes(wfn, x_vec, "cphf_solve")
    Co, Cv, F = _wfn_state(wfn)
    nocc = Co.shape[1]
    Fmo = wfn.Ca.T @ F @ wfn.Ca
    fd = np.diag(Fmo)
    precon = -fd[:nocc, None] + fd[None, nocc:]
    n = len(x_vec)
    ret = [x / precon for x in x_vec]
    r = [x.copy() for x in x_vec]
    Ax = cphf_Hx(wfn, ret)
    nbuild = n
    resid_denom = [max(float(np.sum(x * x)), 1e-14) for x in x_vec]
    rms = [0.0] * n
    for i in range(n):
        r[i] -= Ax[i]
        rms[i] = np.sqrt(float(np.sum(r[i] * r[i])) / resid_denom[i])
    z = [ri / precon for ri in r]
    p = [zi.copy() for zi in z]
    active = [True] * n
    rzpre = [0.0] * n
    if print_lvl:
        print(f"    Guess  mean rms {np.mean(rms):.3e}  max rms {max(rms):.3e}")
    converged = False
    for it in range(1, max_iter):
        idx = [i for i in range(n) if active[i]]
        Ap = cphf_Hx(wfn, [p[i] for i in idx])
        nbuild += len(idx)
        for k, i in enumerate(idx):
            rzpre[i] = float(np.sum(r[i] * z[i]))
            denom = float(np.sum(p[i] * Ap[k]))
            alpha = rzpre[i] / denom if denom != 0.0 else float("nan")
            if np.isnan(alpha):
                active[i] = False
                alpha = 0.0
            ret[i] = ret[i] + alpha * p[i]
            r[i] = r[i] - alpha * Ap[k]
            rms[i] = np.sqrt(float(np.sum(r[i] * r[i])) / resid_denom[i])
            if rms[i] < conv_tol:
                active[i] = False
        if print_lvl:
            print(f"    {it:5d}  mean rms {np.mean(rms):.3e}  max rms {max(rms):.3e}  remain {sum(active)}")
        if max(rms) < conv_tol and not any(active):
            converged = True
            break
        for i in range(n):
            if not active[i]:
                continue
            z[i] = r[i] / precon
            beta = float(np.sum(r[i] * z[i])) / rzpre[i]
            p[i] = beta * p[i] + z[i]
    wfn.cphf_converged = converged or not any(active)
    wfn.cphf_nfock_builds = nbuild
    if raise_on_fail and not wfn.cphf_converged:
        raise CPHFConvergenceError(f"CPHF: {sum(active)} equations did not converge")
    return ret


def cpscf_linear_response(wfn, *props, conv_tol=1e-5, max_iter=10, print_lvl=0):
    """Static response tensors, scf_response.py:55-181.  Supported: "DIPOLE_POLARIZABILITIES".

    The perturbation vectors are psi4's ao_dipole scaled by -1, i.e. the dipole integrals
    <mu|r|nu> (ao_dipole carries the electron charge), transformed to occ x vir; the
    response tensor is -4 <b_i, x_j> for a restricted reference."""
    out = []
    Co, Cv, _ = _wfn_state(wfn)
    for prop in props:
        if not isinstance(prop, str):
            raise TypeError("Property name must be of type string.")
        if prop.upper() != "DIPOLE_POLARIZABILITIES":
            raise ValueError(f"Do not understand '{prop}'.")
        mu = ints.dipole(wfn.basis)          # <mu|r|nu> = psi4's ao_dipole scaled by -1
        vecs = [Co.T @ m @ Cv for m in mu]
        resp = cphf_solve(wfn, vecs, conv_tol, max_iter, print_lvl)
        out.append(np.array([[-4.0 * float(np.sum(vi * xj)) for xj in resp] for vi in vecs]))
    return out
