"""Range-separated exchange (wK) on the GPU against the oracle.

* water / cc-pVDZ, omega = 0.4: the device's attenuated tensor equals L^-1 (Q|w|mn) from
  the host integrals (Cholesky fit, 1e-10); wK for a symmetric and a non-symmetric pair
  equals the oracle's restatement of DFHelper::compute_wK on the device's own tensors
  (1e-11 of the largest element), and K is unchanged by do_wK;
* (H2O)_8 after an omega change (the tensor is rebuilt): wK against the oracle built from
  the host integrals with psi4's J^-1 route (Cholesky metric, 1e-10);
* omega -> infinity: the device wK equals the device K (1e-8);
* combined exchange (wcombine, omega_alpha, omega_beta): wK = alpha K + beta wK(omega) against
  the oracle's restatement of psi4's combined tensors on the device's own B and Bw (1e-11),
  K zero and J unchanged, then a change of (alpha, beta) between computes (tensor rebuilt);
  the symmetric-pair wK is hermitivitized like MemDFJK::compute_JK (exactly symmetric).
"""
import numpy as np
import pytest
from scipy.linalg import solve_triangular

import psi4_amd.molecule as M
from conftest import announce
from oracle.dfhelper import MemDFJKOracle
from psi4_amd import ints
from psi4_amd.basisset import BasisSet

pytestmark = pytest.mark.gpu
OMEGA = 0.4


def _jk(mol, omega):
    from psi4_amd.core import JK
    prim = BasisSet.build(mol, "cc-pvdz")
    aux = BasisSet.build(mol, "cc-pvdz-jkfit")
    jk = JK.build(prim, aux, do_wK=True, memory=4_000_000_000)
    jk.set_print(0)
    jk.set_omega(omega)
    jk.initialize()
    return prim, aux, jk


def _rel(a, b):
    return float(np.abs(a - b).max() / max(1e-300, np.abs(b).max()))


def test_wk_water_vs_oracle(gpu_available):
    mol = M.water_zmat(0.96, 104.5)
    announce("wK water cc-pVDZ omega 0.4: tensor, symmetric and non-symmetric pairs")
    prim, aux, jk = _jk(mol, OMEGA)
    naux, n = jk.info()["naux"], prim.nbf()
    B = jk.get_B_dense(0, naux)
    Bw = jk.get_Bw_dense(0, naux)
    W = ints.three_index(prim, aux, omega=OMEGA)
    ora = MemDFJKOracle(prim, aux, B=B, Bw=Bw)
    W[:, ~ora.sp.fun_mask] = 0.0
    L = np.linalg.cholesky(ints.metric(aux))
    Bw_ref = solve_triangular(L, W.reshape(naux, -1), lower=True).reshape(naux, n, n)
    assert _rel(Bw, Bw_ref) < 1e-10
    rng = np.random.default_rng(11)
    C = np.linalg.qr(rng.standard_normal((n, n)))[0][:, :5]
    jk.C_clear()
    jk.C_left_add(C)
    jk.compute()
    e = _rel(jk.wK()[0], ora.compute_wK([C])[0])
    print(f"  symmetric pair wK rel err {e:.2e}", flush=True)
    assert e < 1e-11
    assert _rel(jk.K()[0], ora.compute([C])[2][0]) < 1e-11
    Cl, Cr = rng.standard_normal((n, 7)), rng.standard_normal((n, 7))
    jk.C_clear()
    jk.C_left_add(Cl)
    jk.C_right_add(Cr)
    jk.compute()
    e = _rel(jk.wK()[0], ora.compute_wK([Cl], [Cr])[0])
    print(f"  non-symmetric pair wK rel err {e:.2e}", flush=True)
    assert e < 1e-11


def test_wk_cluster_omega_change(gpu_available):
    mol = M.water_cluster((2, 2, 2), seed=4)
    announce("wK (H2O)_8 cc-pVDZ, omega 0.4 then 0.25 (attenuated tensor rebuilt)")
    prim, aux, jk = _jk(mol, OMEGA)
    n = prim.nbf()
    rng = np.random.default_rng(12)
    C = np.linalg.qr(rng.standard_normal((n, n)))[0][:, :40]
    jk.C_clear()
    jk.C_left_add(C)
    jk.compute()
    first = jk.wK()[0].copy()
    jk.set_omega(0.25)
    jk.compute()
    ora = MemDFJKOracle(prim, aux, fit="cholesky", omega=0.25)
    e = _rel(jk.wK()[0], ora.compute_wK([C])[0])
    print(f"  (H2O)_8 omega 0.25 wK rel err vs host-integral oracle {e:.2e}", flush=True)
    assert e < 1e-10
    assert _rel(first, jk.wK()[0]) > 1e-3


def test_wk_large_omega_is_K(gpu_available):
    mol = M.water_zmat(0.96, 104.5)
    announce("wK water, omega 1e8 against K")
    prim, aux, jk = _jk(mol, 1e8)
    rng = np.random.default_rng(13)
    C = np.linalg.qr(rng.standard_normal((prim.nbf(), prim.nbf())))[0][:, :5]
    jk.C_clear()
    jk.C_left_add(C)
    jk.compute()
    assert _rel(jk.wK()[0], jk.K()[0]) < 1e-8


def test_wk_combined_exchange(gpu_available):
    mol = M.water_zmat(0.96, 104.5)
    announce("wK water cc-pVDZ omega 0.4: wcombine (alpha 0.75, beta -0.4), then (0.25, 1.0)")
    prim, aux, jk = _jk(mol, OMEGA)
    naux, n = jk.info()["naux"], prim.nbf()
    B = jk.get_B_dense(0, naux)
    Bw = jk.get_Bw_dense(0, naux)
    rng = np.random.default_rng(12)
    C = np.linalg.qr(rng.standard_normal((n, n)))[0][:, :5]
    Cl, Cr = rng.standard_normal((n, 6)), rng.standard_normal((n, 6))
    jk.C_clear()
    jk.C_left_add(Cl)
    jk.C_right_add(Cr)
    jk.compute()
    J0 = jk.J()[0].copy()
    for a, b in ((0.75, -0.4), (0.25, 1.0)):
        jk.set_wcombine(True)
        jk.set_omega_alpha(a)
        jk.set_omega_beta(b)
        ora = MemDFJKOracle(prim, aux, B=B, Bw=Bw, wcombine=True, omega_alpha=a, omega_beta=b)
        jk.C_clear()
        jk.C_left_add(Cl)
        jk.C_right_add(Cr)
        jk.compute()
        e = _rel(jk.wK()[0], ora.compute_wK([Cl], [Cr])[0])
        print(f"  ({a}, {b}) non-symmetric pair combined wK rel err {e:.2e}", flush=True)
        assert e < 1e-11
        assert not jk.K()[0].any()
        assert _rel(jk.J()[0], J0) < 1e-14
        jk.C_clear()
        jk.C_left_add(C)
        jk.compute()
        w = jk.wK()[0]
        assert np.array_equal(w, w.T)
        e = _rel(w, ora.compute_wK([C])[0])
        print(f"  ({a}, {b}) symmetric pair combined wK rel err {e:.2e}", flush=True)
        assert e < 1e-11
    # switching the combination off restores the plain attenuated tensor
    jk.set_wcombine(False)
    assert _rel(jk.get_Bw_dense(0, naux), Bw) < 1e-14
