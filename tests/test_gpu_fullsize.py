"""The synthetic tensor-level workload at its full size (nbf 1500, naux 6000, nocc 300; B is
55 GB in FP64) through properties that do not depend on an oracle run over the whole tensor:

* J and K are additive over the auxiliary index: the builds over Q in [0, 3000) and
  [3000, 6000) sum to the full build (1e-12 of the largest element); K is exactly symmetric;
* a 64-function Q-slice of the device build equals the oracle's compute_J_symm / compute_K
  on the same slice of the device's tensor (1e-11 of the largest element);
* the mixed precisions against FP64 on the same tensor: full K in FP32 within 2e-6 and in
  3xBF16 within 2e-4 of max|K|; the incremental 3xBF16 build of dD (cross form, 600
  columns, C_old a 1e-4 perturbation of C) changes E = <D, 2J - K> by at most 1e-6 Eh
  against FP64 (the per-build target of the mixed path).
"""
import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu


def _build(jk, Cl, Cr=None):
    jk.C_clear()
    jk.C_left_add(Cl)
    if Cr is not None:
        jk.C_right_add(Cr)
    jk.compute()
    return jk.J()[0].copy(), jk.K()[0].copy()


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_full_size_synthetic(gpu_available):
    from oracle.dfhelper import compute_J_symm, compute_K, prepare_sparsity
    from psi4_amd.core import JK, PREC_BF16X3, PREC_FP32, PREC_FP64
    from psi4_amd.workloads import SYNTHETIC, synthetic_C
    nbf, naux, nocc = SYNTHETIC["nbf"], SYNTHETIC["naux"], SYNTHETIC["nocc"]
    announce(f"full-size synthetic nbf {nbf} naux {naux} nocc {nocc}: generate B")
    jk = JK.build_tensor(nbf, naux)
    jk.set_print(0)
    jk.initialize()
    jk.fill_synthetic(SYNTHETIC["seed_B"])
    C = synthetic_C(nbf, nocc, SYNTHETIC["seed_C"])

    announce("full-size synthetic: FP64 build and aux-range additivity")
    J, K = _build(jk, C)
    assert np.array_equal(K, K.T)
    jk.set_aux_range(0, 3000)
    J1, K1 = _build(jk, C)
    jk.set_aux_range(3000, naux)
    J2, K2 = _build(jk, C)
    print(f"  additivity J {_rel(J1 + J2, J):.2e} K {_rel(K1 + K2, K):.2e}", flush=True)
    assert _rel(J1 + J2, J) < 1e-12 and _rel(K1 + K2, K) < 1e-12

    announce("full-size synthetic: Q-slice [0, 64) against the oracle")
    q1 = 64
    jk.set_aux_range(0, q1)
    Js, Ks = _build(jk, C)
    jk.set_aux_range(0, naux)
    Bq = jk.get_B_dense(0, q1)
    sp = prepare_sparsity(np.ones((nbf, nbf)), np.ones((1, 1)), 1.0, 1.0, q1)
    Ppq = np.ascontiguousarray(Bq.transpose(1, 0, 2)).reshape(-1)
    del Bq
    Jo, Ko = np.zeros((nbf, nbf)), np.zeros((nbf, nbf))
    compute_J_symm(Ppq, sp, C @ C.T, Jo)
    compute_K(Ppq, sp, C, C, Ko, True)
    print(f"  slice J {_rel(Js, Jo):.2e} K {_rel(Ks, Ko):.2e}", flush=True)
    assert _rel(Js, Jo) < 1e-11 and _rel(Ks, Ko) < 1e-11

    announce("full-size synthetic: mixed precisions against FP64")
    for prec, tol in ((PREC_FP32, 2e-6), (PREC_BF16X3, 2e-4)):
        jk.set_precision(prec)
        try:
            _, Km = _build(jk, C)
        finally:
            jk.set_precision(PREC_FP64)
        print(f"  precision {prec}: full K rel {_rel(Km, K):.2e}", flush=True)
        assert _rel(Km, K) < tol

    rng = np.random.default_rng(SYNTHETIC["seed_C"] + 1)
    Co = np.linalg.qr(C + 1e-4 * rng.standard_normal(C.shape))[0]
    u, _, vt = np.linalg.svd(Co.T @ C)
    Co = Co @ (u @ vt)
    A, V = 0.5 * (C + Co), C - Co
    Cl, Cr = np.hstack([A, V]), np.hstack([V, A])
    D = C @ C.T
    Jd, Kd = _build(jk, Cl, Cr)
    jk.set_precision(PREC_BF16X3)
    try:
        Jm, Km = _build(jk, Cl, Cr)
    finally:
        jk.set_precision(PREC_FP64)
    dE = abs(float(np.vdot(D, 2.0 * (Jm - Jd) - (Km - Kd))))
    print(f"  incremental 3xBF16: |dE| vs FP64 {dE:.2e} Eh", flush=True)
    assert dE < 1e-6
