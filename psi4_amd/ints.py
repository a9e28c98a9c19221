"""ctypes bindings of the host integral engine (include/psi4amd_ints.h)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._libs import load_ints_lib

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _check(lib, rc):
    if rc != 0:
        raise RuntimeError(lib.p4a_ints_last_error().decode())


def default_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def overlap(basis):
    lib = load_ints_lib()
    n = basis.nbf()
    S = np.zeros((n, n))
    cb, _keep = basis.as_c()
    _check(lib, lib.p4a_ints_overlap(ctypes.byref(cb), _ptr(S)))
    return S


def kinetic(basis):
    lib = load_ints_lib()
    n = basis.nbf()
    T = np.zeros((n, n))
    cb, _keep = basis.as_c()
    _check(lib, lib.p4a_ints_kinetic(ctypes.byref(cb), _ptr(T)))
    return T


def dipole(basis, origin=(0.0, 0.0, 0.0)):
    """<mu| r - origin |nu> as [3, nbf, nbf] (x, y, z; no electron charge)."""
    lib = load_ints_lib()
    n = basis.nbf()
    M = np.zeros((3, n, n))
    cb, _keep = basis.as_c()
    o = np.ascontiguousarray(origin, np.float64)
    _check(lib, lib.p4a_ints_dipole(ctypes.byref(cb), _ptr(o), _ptr(M)))
    return M


def angular_momentum(basis, origin=(0.0, 0.0, 0.0)):
    """<mu| (r - origin) x nabla |nu> as [3, nbf, nbf] (antisymmetric; no factor -i)."""
    lib = load_ints_lib()
    n = basis.nbf()
    M = np.zeros((3, n, n))
    cb, _keep = basis.as_c()
    o = np.ascontiguousarray(origin, np.float64)
    _check(lib, lib.p4a_ints_angular_momentum(ctypes.byref(cb), _ptr(o), _ptr(M)))
    return M


def potential(basis, mol):
    lib = load_ints_lib()
    n = basis.nbf()
    V = np.zeros((n, n))
    cb, _keep = basis.as_c()
    Z = np.ascontiguousarray(mol.Z, np.float64)
    xyz = np.ascontiguousarray(mol.xyz, np.float64)
    _check(lib, lib.p4a_ints_potential(ctypes.byref(cb), mol.natom, _ptr(Z), _ptr(xyz), _ptr(V)))
    return V


def metric(aux, nthreads=None, omega=None):
    """(P|Q); with omega > 0 over erf(omega r)/r."""
    lib = load_ints_lib()
    n = aux.nbf()
    J = np.zeros((n, n))
    cb, _keep = aux.as_c()
    if omega:
        _check(lib, lib.p4a_ints_metric_erf(ctypes.byref(cb), ctypes.c_double(omega), _ptr(J),
                                            nthreads or default_threads()))
    else:
        _check(lib, lib.p4a_ints_metric(ctypes.byref(cb), _ptr(J), nthreads or default_threads()))
    return J


def three_index(prim, aux, nthreads=None, omega=None):
    """Dense (P|mn) as an array [naux, nbf, nbf]; with omega > 0 over erf(omega r)/r."""
    lib = load_ints_lib()
    n, na = prim.nbf(), aux.nbf()
    out = np.zeros((na, n, n))
    dst = np.arange(n * n, dtype=np.int64)
    cp, _k1 = prim.as_c()
    ca, _k2 = aux.as_c()
    dptr = dst.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    if omega:
        _check(lib, lib.p4a_ints_three_index_erf(ctypes.byref(cp), ctypes.byref(ca), ctypes.c_double(omega), dptr,
                                                 ctypes.c_int64(n * n), None, _ptr(out), nthreads or default_threads()))
    else:
        _check(lib, lib.p4a_ints_three_index(ctypes.byref(cp), ctypes.byref(ca), dptr, ctypes.c_int64(n * n), None,
                                             _ptr(out), nthreads or default_threads()))
    return out


def schwarz(prim, nthreads=None):
    """(fun_max [nbf,nbf], shell_max [nshell,nshell], max_val) per dfhelper.cc:322-366."""
    lib = load_ints_lib()
    n, ns = prim.nbf(), prim.nshell
    fm = np.zeros((n, n))
    sm = np.zeros((ns, ns))
    mx = ctypes.c_double(0)
    cp, _k = prim.as_c()
    _check(lib, lib.p4a_ints_schwarz(ctypes.byref(cp), _ptr(fm), _ptr(sm), ctypes.byref(mx),
                                     nthreads or default_threads()))
    return fm, sm, mx.value


def four_index(basis, nthreads=None, omega=None):
    """Full (mn|ls); with omega > 0 over erf(omega r)/r."""
    lib = load_ints_lib()
    n = basis.nbf()
    out = np.zeros((n, n, n, n))
    cb, _k = basis.as_c()
    if omega:
        _check(lib, lib.p4a_ints_four_index_erf(ctypes.byref(cb), ctypes.c_double(omega), _ptr(out),
                                                nthreads or default_threads()))
    else:
        _check(lib, lib.p4a_ints_four_index(ctypes.byref(cb), _ptr(out), nthreads or default_threads()))
    return out


def boys(nmax, T):
    lib = load_ints_lib()
    F = np.zeros(nmax + 1)
    _check(lib, lib.p4a_ints_boys(nmax, ctypes.c_double(T), _ptr(F)))
    return F


# ---- gradient contractions (derivatives with respect to shell centres) ----
def _grad_out(basis):
    return np.zeros((basis.nshell, 3))


def overlap_grad(basis, W):
    """sum_mn W_mn dS_mn/dR per shell centre, [nshell, 3] (W symmetric)."""
    lib = load_ints_lib()
    g = _grad_out(basis)
    cb, _keep = basis.as_c()
    _check(lib, lib.p4a_ints_overlap_grad(ctypes.byref(cb), _ptr(np.ascontiguousarray(W, np.float64)), _ptr(g)))
    return g


def kinetic_grad(basis, D):
    lib = load_ints_lib()
    g = _grad_out(basis)
    cb, _keep = basis.as_c()
    _check(lib, lib.p4a_ints_kinetic_grad(ctypes.byref(cb), _ptr(np.ascontiguousarray(D, np.float64)), _ptr(g)))
    return g


def potential_grad(basis, mol, D):
    """(per-shell-centre part [nshell, 3], per-nucleus part [natom, 3]) of sum_mn D_mn dV_mn/dR."""
    lib = load_ints_lib()
    g = _grad_out(basis)
    gn = np.zeros((mol.natom, 3))
    cb, _keep = basis.as_c()
    Z = np.ascontiguousarray(mol.Z, np.float64)
    xyz = np.ascontiguousarray(mol.xyz, np.float64)
    _check(lib, lib.p4a_ints_potential_grad(ctypes.byref(cb), _ptr(np.ascontiguousarray(D, np.float64)), mol.natom,
                                            _ptr(Z), _ptr(xyz), _ptr(g), _ptr(gn)))
    return g, gn


def three_index_grad(prim, aux, G, nthreads=None):
    """sum_{P,mn} G[P,m,n] d(P|mn)/dR: (per primitive shell [nshell, 3], per aux shell [naux_shell, 3])."""
    lib = load_ints_lib()
    gp, ga = _grad_out(prim), _grad_out(aux)
    cp, _k1 = prim.as_c()
    ca, _k2 = aux.as_c()
    G = np.ascontiguousarray(G, np.float64)
    _check(lib, lib.p4a_ints_three_index_grad(ctypes.byref(cp), ctypes.byref(ca), _ptr(G), _ptr(gp), _ptr(ga),
                                              nthreads or default_threads()))
    return gp, ga


def metric_grad(aux, W, nthreads=None):
    """sum_PQ W_PQ d(P|Q)/dR per aux shell centre, [naux_shell, 3] (W symmetric)."""
    lib = load_ints_lib()
    ga = _grad_out(aux)
    ca, _k = aux.as_c()
    _check(lib, lib.p4a_ints_metric_grad(ctypes.byref(ca), _ptr(np.ascontiguousarray(W, np.float64)), _ptr(ga),
                                         nthreads or default_threads()))
    return ga
