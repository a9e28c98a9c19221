"""GPU DF-JK through the C ABI vs the CPU oracle (restated MemDFJK).

FP64 J and K must match the oracle to 1e-11 max-abs (the bar for plain FP64).
Mixed-precision K is checked against the FP64 oracle with tolerances relative
to max|K|: FP32 MFMA 2e-6, 3-term BF16 split 2e-4.  In both mixed modes J
accumulates in FP64 from the FP32 copy of B (1e-7 of max|J|).
"""
import numpy as np
import pytest

from conftest import announce
from helpers import gpu_jk, maxabs, random_orbitals, water_system

pytestmark = pytest.mark.gpu

TOL = 1e-11
REL_K = {1: 2e-6, 2: 2e-4}  # PREC_FP32, PREC_BF16X3
# J accumulates in FP64 but reads the FP32 copy of B in the mixed modes
REL_J = {1: 1e-7, 2: 1e-7}


@pytest.fixture(scope="module")
def h2o4(gpu_available):
    from oracle.dfhelper import MemDFJKOracle
    mol, prim, aux = water_system((2, 2, 1), seed=0)
    announce("(H2O)_4 cc-pVDZ/cc-pVDZ-jkfit: oracle + GPU initialize")
    ora = MemDFJKOracle(prim, aux)
    jk = gpu_jk(prim, aux)
    return mol, prim, aux, ora, jk


@pytest.fixture(scope="module")
def h2o8(gpu_available):
    from oracle.dfhelper import MemDFJKOracle
    mol, prim, aux = water_system((2, 2, 2), seed=3)
    announce("(H2O)_8 cc-pVDZ/cc-pVDZ-jkfit: oracle + GPU initialize")
    ora = MemDFJKOracle(prim, aux)
    jk = gpu_jk(prim, aux)
    return mol, prim, aux, ora, jk


def _run(jk, lefts, rights=None):
    jk.C_clear()
    for L in lefts:
        jk.C_left_add(L)
    for R in rights or []:
        jk.C_right_add(R)
    jk.compute()
    return jk.D(), jk.J(), jk.K()


def test_symmetric_jk_h2o4(h2o4):
    mol, prim, aux, ora, jk = h2o4
    C = random_orbitals(prim.nbf(), 20, seed=1)
    announce("symmetric J/K nocc=20")
    D, J, K = _run(jk, [C])
    Do, Jo, Ko = ora.compute([C])
    assert maxabs(D[0], Do[0]) < 1e-13
    assert maxabs(J[0], Jo[0]) < TOL
    assert maxabs(K[0], Ko[0]) < TOL


@pytest.mark.parametrize("width", [16, 70, 130, 300, 330])
def test_symmetric_panel_widths_h2o4(h2o4, width):
    mol, prim, aux, ora, jk = h2o4
    C = random_orbitals(prim.nbf(), width, seed=width)
    announce(f"symmetric J/K width={width}")
    D, J, K = _run(jk, [C])
    Do, Jo, Ko = ora.compute([C])
    assert maxabs(J[0], Jo[0]) < TOL * max(1.0, width / 20)
    assert maxabs(K[0], Ko[0]) < TOL * max(1.0, width / 20)


@pytest.mark.parametrize("width", [16, 70, 130, 250, 300, 330])
def test_bf16x3_panel_widths_h2o4(h2o4, width):
    """The BF16x3 kernels at every panel width (64-column blocks per wave 1-5, one and two
    panels: the direct loads of the C panel are dealt to the waves by width) against the
    oracle, relative to max|K| and max|J|."""
    mol, prim, aux, ora, jk = h2o4
    C = random_orbitals(prim.nbf(), width, seed=width)
    announce(f"symmetric BF16x3 J/K width={width}")
    jk.set_precision(2)
    try:
        _, J, K = _run(jk, [C])
    finally:
        jk.set_precision(0)
    _, Jo, Ko = ora.compute([C])
    assert maxabs(K[0], Ko[0]) < 5 * REL_K[2] * np.abs(Ko[0]).max()
    assert maxabs(J[0], Jo[0]) < 5 * REL_J[2] * np.abs(Jo[0]).max()


def test_nonsymmetric_multi_pair_h2o4(h2o4):
    mol, prim, aux, ora, jk = h2o4
    nbf = prim.nbf()
    pairs = [(random_orbitals(nbf, w, seed=10 + w), random_orbitals(nbf, w, seed=100 + w)) for w in (16, 23, 30)]
    announce("non-symmetric J/K, 3 pairs of widths 16/23/30")
    D, J, K = _run(jk, [p[0] for p in pairs], [p[1] for p in pairs])
    Do, Jo, Ko = ora.compute([p[0] for p in pairs], [p[1] for p in pairs])
    for i in range(3):
        assert maxabs(D[i], Do[i]) < 1e-13
        assert maxabs(J[i], Jo[i]) < TOL
        assert maxabs(K[i], Ko[i]) < TOL


def test_incremental_signed_pair_h2o8(h2o8):
    """C_left=[C_new,C_old], C_right=[C_new,-C_old] gives J(dD), K(dD) with dD = D_new - D_old."""
    mol, prim, aux, ora, jk = h2o8
    nbf = prim.nbf()
    Cn = random_orbitals(nbf, 40, seed=5)
    Co = Cn + 1e-3 * random_orbitals(nbf, 40, seed=6)
    L = np.hstack([Cn, Co])
    R = np.hstack([Cn, -Co])
    announce("signed incremental pair (H2O)_8 width 80")
    D, J, K = _run(jk, [L], [R])
    Do, Jo, Ko = ora.compute([L], [R])
    assert maxabs(D[0], Do[0]) < 1e-13
    assert maxabs(J[0], Jo[0]) < TOL
    assert maxabs(K[0], Ko[0]) < TOL
    # linearity: K(dD) = K(D_new) - K(D_old)
    _, Jn, Kn = ora.compute([Cn])
    _, Jold, Kold = ora.compute([Co])
    assert maxabs(K[0], Kn[0] - Kold[0]) < TOL
    assert maxabs(J[0], Jn[0] - Jold[0]) < TOL


def test_multi_block_geometry_h2o8(h2o8):
    mol, prim, aux, ora, jk = h2o8
    C = random_orbitals(prim.nbf(), 40, seed=9)
    announce("(H2O)_8 symmetric width 40 (several 128-row blocks)")
    D, J, K = _run(jk, [C])
    Do, Jo, Ko = ora.compute([C])
    assert maxabs(J[0], Jo[0]) < TOL
    assert maxabs(K[0], Ko[0]) < TOL


@pytest.mark.parametrize("prec", [1, 2])
def test_mixed_precision_symmetric_h2o8(h2o8, prec):
    mol, prim, aux, ora, jk = h2o8
    C = random_orbitals(prim.nbf(), 40, seed=11)
    announce(f"mixed precision {prec} symmetric (H2O)_8")
    jk.set_precision(prec)
    try:
        D, J, K = _run(jk, [C])
    finally:
        jk.set_precision(0)
    Do, Jo, Ko = ora.compute([C])
    assert maxabs(J[0], Jo[0]) < REL_J[prec] * np.abs(Jo[0]).max()
    scale = np.abs(Ko[0]).max()
    assert maxabs(K[0], Ko[0]) < REL_K[prec] * scale
    assert maxabs(K[0], K[0].T) < 1e-14 * scale  # symmetric by construction


@pytest.mark.parametrize("prec", [1, 2])
def test_mixed_precision_incremental_h2o8(h2o8, prec):
    """Mixed-precision K(dD) in the stable incremental form
    dD = C_new dC^T + dC C_old^T (C_left=[C_new, dC], C_right=[dC, C_old], dC = C_new - C_old):
    every product carries an O(|dC|) factor, so the error scales with |dD|."""
    mol, prim, aux, ora, jk = h2o8
    nbf = prim.nbf()
    Cn = random_orbitals(nbf, 40, seed=21)
    dC = 1e-4 * random_orbitals(nbf, 40, seed=22)
    Co = Cn - dC
    announce(f"mixed precision {prec} stable incremental (H2O)_8")
    jk.set_precision(prec)
    try:
        D, J, K = _run(jk, [np.hstack([Cn, dC])], [np.hstack([dC, Co])])
    finally:
        jk.set_precision(0)
    _, Jn, Kn = ora.compute([Cn])
    _, Jold, Kold = ora.compute([Co])
    ref = Kn[0] - Kold[0]
    dJ = Jn[0] - Jold[0]
    assert maxabs(J[0], dJ) < 5 * REL_J[prec] * np.abs(dJ).max()
    # error relative to |dK| stays at the precision's relative level
    assert maxabs(K[0], ref) < 5 * REL_K[prec] * np.abs(ref).max()


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_cross_incremental_h2o8(h2o8, prec):
    """Cross form dD = A V^T + V A^T (C_left=[A, V], C_right=[V, A], A = (C_new + C_old)/2,
    V = C_new - C_old): FP64 to 1e-11 against K(C_new) - K(C_old); mixed precision relative
    to |dK|; K(dD) exactly symmetric at every precision."""
    mol, prim, aux, ora, jk = h2o8
    nbf = prim.nbf()
    Cn = random_orbitals(nbf, 40, seed=41)
    V = 1e-4 * random_orbitals(nbf, 40, seed=42)
    Co = Cn - V
    A = 0.5 * (Cn + Co)
    announce(f"cross incremental form, precision {prec}, (H2O)_8")
    jk.set_precision(prec)
    try:
        D, J, K = _run(jk, [np.hstack([A, V])], [np.hstack([V, A])])
    finally:
        jk.set_precision(0)
    _, Jn, Kn = ora.compute([Cn])
    _, Jold, Kold = ora.compute([Co])
    ref = Kn[0] - Kold[0]
    assert maxabs(D[0], Cn @ Cn.T - Co @ Co.T) < 1e-14
    assert maxabs(K[0], K[0].T) == 0.0
    if prec == 0:
        assert maxabs(K[0], ref) < TOL and maxabs(J[0], Jn[0] - Jold[0]) < TOL
    else:
        assert maxabs(K[0], ref) < 5 * REL_K[prec] * np.abs(ref).max()
        dJ = Jn[0] - Jold[0]
        assert maxabs(J[0], dJ) < 5 * REL_J[prec] * np.abs(dJ).max()


def test_empty_pair_beside_a_full_one_h2o4(h2o4):
    """A pair with no columns (a spin with no occupied orbitals) next to a full one: psi4 gives
    D = 0, J = 0 and skips K for nocc == 0 (dfhelper.cc:3355-3358), and the full pair is
    unaffected; the same with a single column."""
    mol, prim, aux, ora, jk = h2o4
    nbf = prim.nbf()
    C = random_orbitals(nbf, 20, seed=5)
    c1 = random_orbitals(nbf, 1, seed=6)
    announce("empty and one-column pairs beside a full pair")
    D, J, K = _run(jk, [C, np.zeros((nbf, 0)), c1])
    Do, Jo, Ko = ora.compute([C, c1])
    assert maxabs(J[0], Jo[0]) < TOL and maxabs(K[0], Ko[0]) < TOL
    assert np.abs(D[1]).max() == 0.0 and np.abs(J[1]).max() == 0.0 and np.abs(K[1]).max() == 0.0
    assert maxabs(J[2], Jo[1]) < TOL and maxabs(K[2], Ko[1]) < TOL


def test_accumulate_and_aux_range_h2o4(h2o4):
    mol, prim, aux, ora, jk = h2o4
    naux = aux.nbf()
    C = random_orbitals(prim.nbf(), 20, seed=31)
    announce("aux-range additivity and accumulate mode")
    _, Jfull, Kfull = _run(jk, [C])
    Jfull, Kfull = Jfull[0].copy(), Kfull[0].copy()
    jk.set_aux_range(0, naux // 3)
    _, J1, K1 = _run(jk, [C])
    J1, K1 = J1[0].copy(), K1[0].copy()
    jk.set_aux_range(naux // 3, naux)
    jk.set_accumulate(True)
    try:
        jk.compute()   # same C list; adds into the previous outputs
        _, J2, K2 = jk.D(), jk.J(), jk.K()
    finally:
        jk.set_accumulate(False)
        jk.set_aux_range(0, naux)
    assert maxabs(J2[0], Jfull) < TOL
    assert maxabs(K2[0], Kfull) < TOL


def test_synthetic_generator_matches_oracle(gpu_available):
    from oracle.synthetic import synthetic_B
    from psi4_amd.core import JK
    nbf, naux, seed = 70, 9, 424242
    announce("synthetic B generator nbf=70 naux=9")
    jk = JK.build_tensor(nbf, naux)
    jk.set_print(0)
    jk.initialize()
    jk.fill_synthetic(seed)
    Bg = jk.get_B_dense(0, naux)
    Bo = synthetic_B(nbf, 0, naux, seed)
    assert np.abs(Bg - Bo).max() < 1e-14 * np.abs(Bo).max()
    assert np.abs(Bg - Bg.transpose(0, 2, 1)).max() == 0.0


@pytest.mark.parametrize("shape", [(420, 96, 200), (1500, 128, 300)])
@pytest.mark.parametrize("prec", [1, 2])
def test_mixed_two_panel_incremental_synthetic(gpu_available, prec, shape):
    """Stable incremental form at ncol = 2*nocc > 320 (two column panels) on a synthetic tensor:
    mixed K(dD) vs FP64 K(dD) on identical inputs, error relative to |K(dD)|."""
    from psi4_amd.core import JK
    from psi4_amd.workloads import synthetic_C
    nbf, naux, nocc = shape
    jk = JK.build_tensor(nbf, naux)
    jk.set_print(0)
    jk.initialize()
    jk.fill_synthetic(99)
    C = synthetic_C(nbf, nocc, 5)
    rng = np.random.default_rng(6)
    Co = np.linalg.qr(C + 1e-4 * rng.standard_normal(C.shape))[0]
    u, _, vt = np.linalg.svd(Co.T @ C)
    Co = Co @ (u @ vt)
    dC = C - Co
    L, R = np.hstack([C, dC]), np.hstack([dC, Co])
    out = {}
    for p in (0, prec):
        announce(f"two-panel stable incremental, precision {p}")
        jk.set_precision(p)
        try:
            _, J, K = _run(jk, [L], [R])
        finally:
            jk.set_precision(0)
        out[p] = (J[0].copy(), K[0].copy())
        # full symmetric build at the same precision for scale
    _, _, Kfull = _run(jk, [C])
    err = np.abs(out[prec][1] - out[0][1]).max()
    scale_d = np.abs(out[0][1]).max()
    D = C @ C.T
    dE = abs(np.vdot(D, out[prec][1] - out[0][1]))
    print(f"  prec {prec} shape {shape}: |dK| max {scale_d:.3e}  err {err:.3e}  rel {err / scale_d:.3e}  "
          f"|K| max {np.abs(Kfull[0]).max():.3e}  |tr(D dK_err)| {dE:.3e}", flush=True)
    assert scale_d > 1e-6 * np.abs(Kfull[0]).max()
    assert np.abs(out[prec][0] - out[0][0]).max() < 5 * REL_J[prec] * np.abs(out[0][0]).max()
    assert err < 5 * REL_K[prec] * scale_d


@pytest.mark.parametrize("var,val", [("PSI4AMD_HT_ROWWAVES", "1"), ("PSI4AMD_HT_ROWWAVES", "3"),
                                     ("PSI4AMD_HT_ROWWAVES", "4"), ("PSI4AMD_KGEMM_DC", "0")])
def test_kernel_shapes_agree_bitwise_synthetic(gpu_available, var, val):
    """The BF16x3 half-transform's other workgroup shapes (PSI4AMD_HT_ROWWAVES: 1 = 4 waves
    over 64 rows, 3 = two 4-wave 128x(32*nbw) workgroups per CU, 4 = C staged through
    registers) and the contraction with its operands staged through registers
    (PSI4AMD_KGEMM_DC=0) run the same MFMAs in the same order for every element, so J and K
    equal the defaults' bit for bit (two column panels of 256: nbf 420, 2*nocc = 400)."""
    import os
    from psi4_amd.core import JK
    from psi4_amd.workloads import synthetic_C
    nbf, naux, nocc = 420, 96, 200
    default = {"PSI4AMD_HT_ROWWAVES": "0", "PSI4AMD_KGEMM_DC": "1"}[var]
    C = synthetic_C(nbf, nocc, 5)
    rng = np.random.default_rng(6)
    Co = np.linalg.qr(C + 1e-4 * rng.standard_normal(C.shape))[0]
    u, _, vt = np.linalg.svd(Co.T @ C)
    Co = Co @ (u @ vt)
    A, V = 0.5 * (C + Co), C - Co

    def run(v):
        # the setting is read when a handle is initialised and applies to every handle
        os.environ[var] = v
        jk = JK.build_tensor(nbf, naux)
        jk.set_print(0)
        jk.initialize()
        jk.fill_synthetic(99)
        jk.set_precision(2)
        _, J, K = _run(jk, [np.hstack([A, V])], [np.hstack([V, A])])
        return J[0].copy(), K[0].copy()

    before = os.environ.get(var)
    try:
        announce(f"{var}={val} vs the default, BF16x3 cross form")
        J0, K0 = run(default)
        J1, K1 = run(val)
    finally:
        run(default)  # restores the default for the handles of later tests
        if before is None:
            del os.environ[var]
        else:
            os.environ[var] = before
    assert np.abs(K0).max() > 0
    assert np.array_equal(K1, K0)
    assert np.array_equal(J1, J0)
