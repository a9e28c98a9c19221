"""Out-of-core B (SCF_SUBTYPE OUT_OF_CORE; AUTO only when B does not fit the device): the fitted
tensor stays in pinned host memory and each compute streams it through the device in
Q-blocks.  On (H2O)_8 / cc-pVDZ with a memory budget that forces ~20 blocks, J and K equal
the in-core handle's (same kernels, blocked sums: 1e-13 of the largest element) for a
symmetric pair, a general non-symmetric pair and the mixed-precision cross form, and the
streamed B read back equals the in-core B exactly.  The energy adjoint's E and dE/dC equal the
in-core ones out of core; its dE/dB needs B in core and is refused.
"""
import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu


def _jk(mol, subalgo=None, memory=None):
    """Both handles take (P|mn) from the host engine (out of core always does), so their
    tensors are bit-identical and J/K differ only by the blocking."""
    import os
    from psi4_amd.basisset import BasisSet
    from psi4_amd.core import JK
    prim = BasisSet.build(mol, "cc-pvdz")
    aux = BasisSet.build(mol, "cc-pvdz-jkfit")
    jk = JK.build(prim, aux, do_wK=False, memory=memory or 2_000_000_000)
    jk.set_print(0)
    if subalgo:
        jk.set_subalgo(subalgo)
    old = os.environ.get("PSI4AMD_HOST_INTS")
    os.environ["PSI4AMD_HOST_INTS"] = "1"
    try:
        jk.initialize()
    finally:
        if old is None:
            os.environ.pop("PSI4AMD_HOST_INTS")
        else:
            os.environ["PSI4AMD_HOST_INTS"] = old
    return jk


def _run(jk, Cl, Cr=None, prec=None):
    from psi4_amd.core import PREC_BF16X3, PREC_FP64
    jk.set_precision(PREC_BF16X3 if prec == "bf16x3" else PREC_FP64)
    jk.C_clear()
    jk.C_left_add(Cl)
    if Cr is not None:
        jk.C_right_add(Cr)
    jk.compute()
    return jk.J()[0].copy(), jk.K()[0].copy()


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_out_of_core_matches_in_core(gpu_available):
    from psi4_amd.molecule import water_cluster
    mol = water_cluster((2, 2, 2), seed=7)
    announce("(H2O)_8 in core")
    core = _jk(mol, "INCORE")
    announce("(H2O)_8 out of core, ~20 Q-blocks from a small memory budget")
    naux = core.info()["naux"]
    ooc = _jk(mol, "OUT_OF_CORE", memory=5_000_000)
    assert ooc.info()["out_of_core"] and not core.info()["out_of_core"]
    # AUTO keeps B in HBM whenever it fits, whatever the (host-sized) memory setting
    assert not _jk(mol, "AUTO", memory=5_000_000).info()["out_of_core"]
    assert "Host-stream" in ooc.print_header()
    assert np.array_equal(ooc.get_B_dense(0, naux), core.get_B_dense(0, naux))
    n = core.info()["nbf"]
    rng = np.random.default_rng(21)
    C = np.linalg.qr(rng.standard_normal((n, n)))[0][:, :40]
    Cl, Cr = rng.standard_normal((n, 13)), rng.standard_normal((n, 13))
    for label, args in (("symmetric", (C,)), ("general", (Cl, Cr))):
        announce(f"(H2O)_8 {label} pair, in core vs out of core")
        Jc, Kc = _run(core, *args)
        Jo, Ko = _run(ooc, *args)
        print(f"  {label}: J {_rel(Jo, Jc):.1e} K {_rel(Ko, Kc):.1e}", flush=True)
        assert _rel(Jo, Jc) < 1e-13 and _rel(Ko, Kc) < 1e-13
    # mixed precision, cross form of a density change
    Cold = np.linalg.qr(C + 1e-3 * rng.standard_normal(C.shape))[0]
    A, V = 0.5 * (C + Cold), C - Cold
    announce("(H2O)_8 bf16x3 cross form, in core vs out of core")
    Jc, Kc = _run(core, np.hstack([A, V]), np.hstack([V, A]), "bf16x3")
    Jo, Ko = _run(ooc, np.hstack([A, V]), np.hstack([V, A]), "bf16x3")
    print(f"  bf16x3 cross: J {_rel(Jo, Jc):.1e} K {_rel(Ko, Kc):.1e}", flush=True)
    assert _rel(Jo, Jc) < 1e-12 and _rel(Ko, Kc) < 1e-6
    # the energy adjoint: E and dE/dC come from J and K, so they work out of core; dE/dB
    # (the size of B) needs B in core
    announce("(H2O)_8 energy adjoint E and dE/dC, in core vs out of core")
    Ec, dCc = core.energy_adjoint(C)[:2]
    Eo, dCo = ooc.energy_adjoint(C)[:2]
    print(f"  adjoint: E {abs(Eo - Ec):.1e}  dE/dC {_rel(dCo, dCc):.1e}", flush=True)
    assert abs(Eo - Ec) < 1e-12 * abs(Ec) and _rel(dCo, dCc) < 1e-12
    with pytest.raises(RuntimeError, match="in core"):
        ooc.energy_adjoint(C, True)


def test_out_of_core_rejects_device_tensor_paths(gpu_available):
    """B streamed from host memory: wK (psi4 refuses it out of core, dfhelper.cc:201-205) and
    the synthetic fill, which both need B on the device, fail with a status before any
    kernel is launched; J/K still run afterwards."""
    from psi4_amd.molecule import water_cluster
    mol = water_cluster((2, 2, 1), seed=0)
    announce("(H2O)_4 out of core: wK and the synthetic fill are refused")
    jk = _jk(mol, subalgo="OUT_OF_CORE")
    assert jk.info()["out_of_core"]
    with pytest.raises(RuntimeError, match="out of core"):
        jk.set_do_wK(True)
    with pytest.raises(RuntimeError, match="resident on the device"):
        jk.fill_synthetic(3)
    C = np.random.default_rng(2).standard_normal((96, 20)) / 10.0
    J, K = _run(jk, C)
    assert np.isfinite(J).all() and np.isfinite(K).all()
