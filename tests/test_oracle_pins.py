"""Pins of the oracle (CPU, no GPU): psi4's stored values, closed forms,
brute force and invariants.  Each check is chosen so that a plausible mistake
(dropped term, wrong sign or index, transposed operand) fails one of them."""
import math

import numpy as np
import pytest
from scipy.special import erf

from oracle.dfhelper import (MemDFJKOracle, build_JK, build_ppq, metric_power, ppq_to_dense,
                             prepare_sparsity)
from oracle.rhf import rhf
from oracle.synthetic import philox4x32_10
from psi4_amd import ints
from psi4_amd.basisset import BasisSet
import psi4_amd.molecule as M

BOHR_2006 = 0.52917720859   # CODATA 2006, the constant psi4 used for the 2017-era reference outputs


def water(r, theta, bohr=BOHR_2006):
    old = M.BOHR2ANGSTROM
    M.BOHR2ANGSTROM = bohr
    try:
        return M.water_zmat(r, theta)
    finally:
        M.BOHR2ANGSTROM = old


def oracle_energy(mol, basis="cc-pvdz", aux="cc-pvdz-jkfit", **kw):
    prim, ab = BasisSet.build(mol, basis), BasisSet.build(mol, aux)
    o = MemDFJKOracle(prim, ab, **kw)
    S = ints.overlap(prim)
    H = ints.kinetic(prim) + ints.potential(prim, mol)

    def jk(C):
        _, J, K = o.compute([C])
        return J[0], K[0]
    return rhf(H, S, mol.nelectron() // 2, mol.nuclear_repulsion_energy(), jk), o


# ------------------------------------------------------------------ psi4 golden values
def test_psi4_dfomp2_1_energy():
    """tests/dfomp2-1/input.dat:3-4: E_nuc 9.18738642147759, E(DF-RHF) -76.02674017978704."""
    xyz = np.array([[0, 0, -0.065655108074], [0, -0.757365949175, 0.520997104936],
                    [0, 0.757365949175, 0.520997104936]])
    mol = M.Molecule(["O", "H", "H"], xyz / BOHR_2006)
    assert abs(mol.nuclear_repulsion_energy() - 9.18738642147759) < 1e-11
    r, _ = oracle_energy(mol)
    assert abs(r["E"] - (-76.02674017978704)) < 1e-9


def test_psi4_scf_guess_read1_energy():
    """tests/scf-guess-read1/input.dat:4-14 (O; H 1 0.96; H 1 0.96 2 104.5): -76.02663273485877."""
    mol = water(0.96, 104.5)
    assert abs(mol.nuclear_repulsion_energy() - 9.168193260387573) < 1e-11
    r, _ = oracle_energy(mol)
    assert abs(r["E"] - (-76.02663273485877)) < 1e-9   # the reference ran to d_convergence 1e-6


def test_psi4_scf2_cc_pvtz_energy():
    """tests/scf2: DF-RHF water / cc-pVTZ (f functions; cc-pVTZ-jkfit up to g),
    E = -76.05098620307962 checked by psi4 to 1e-9 (input.dat:4, 36).  psi4's oriented
    geometry (output.ref:85-87, Angstrom) with the CODATA-2006 Bohr radius of that run
    (E_nuc 8.801465529972067, output.ref:93)."""
    ang = np.array([[0.0, 0.0, -0.068516219310], [0.0, -0.790689573744, 0.543701060724],
                    [0.0, 0.790689573744, 0.543701060724]])
    mol = M.Molecule(["O", "H", "H"], ang / BOHR_2006)
    assert abs(mol.nuclear_repulsion_energy() - 8.801465529972067) < 1e-11
    res, o = oracle_energy(mol, "cc-pvtz", "cc-pvtz-jkfit")
    assert o.nbf == 58 and o.naux == 139
    assert abs(res["E"] - (-76.05098620307962)) < 1e-9


def test_psi4_jkmemory_ar5():
    """tests/pytests/test_jkmemory.py:37: MEM_DF estimate 1,590,520 doubles for Ar5/cc-pVDZ (-n1)."""
    mol = M.Molecule.from_string("Ar 0 0 0\nAr 0 0 5\nAr 0 0 15\nAr 0 0 25\nAr 0 0 35")
    prim, aux = BasisSet.build(mol, "cc-pvdz"), BasisSet.build(mol, "cc-pvdz-jkfit")
    fm, sm, mx = ints.schwarz(prim)
    sp = prepare_sparsity(fm, sm, mx, 1e-12, aux.nbf())
    from oracle.dfhelper import memory_estimate
    assert sp.nsig_pairs() == 1962
    assert memory_estimate(sp, aux.max_nfunc(), nthreads=1) == 1590520


# ------------------------------------------------------------------ closed forms
def test_psi4_ao_angular_momentum_h2o():
    """Element-wise pin of the AO conventions (normalisation, Cartesian and solid-harmonic
    order, s/p/d) of the host integral engine: psi4's stored MintsHelper::ao_angular_momentum
    of H2O / cc-pVDZ (tests/pytests/oei_reference_data.json, compared there at atol 1e-14).
    psi4's matrices are -<mu|(r x nabla)|nu>."""
    import json
    import os
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "psi4_ao_angular_momentum_h2o_ccpvdz.json")))
    mol = M.Molecule([a[0] for a in fx["geometry_bohr"]], np.array([a[1:] for a in fx["geometry_bohr"]]))
    bs = BasisSet.build(mol, fx["basis"])
    ref = np.array(fx["L"]).reshape(fx["L_shape"])
    L = ints.angular_momentum(bs)
    assert np.abs(-L - ref).max() < 1e-14
    assert np.abs(ref).max() > 1.0


@pytest.mark.parametrize("T", [0.0, 1e-8, 0.3, 2.5, 17.0, 44.9, 45.1, 120.0])
def test_boys_F0_closed_form(T):
    F = ints.boys(12, T)
    f0 = 1.0 if T == 0 else 0.5 * math.sqrt(math.pi / T) * erf(math.sqrt(T))
    assert abs(F[0] - f0) < 1e-14 * max(1.0, f0)
    # F_n(0) = 1/(2n+1); downward recursion consistency elsewhere
    if T == 0.0:
        assert np.allclose(F, 1.0 / (2 * np.arange(13) + 1), rtol=1e-14, atol=0)
    else:
        for n in range(12):
            assert abs(F[n] - (2 * T * F[n + 1] + math.exp(-T)) / (2 * n + 1)) < 1e-14 * F[n]


def _s_basis(exps_centers):
    b = BasisSet("s", np.zeros(len(exps_centers), np.int32), np.ones(len(exps_centers), np.int32),
                 np.array([c for _, c in exps_centers], float), np.array([e for e, _ in exps_centers], float),
                 np.ones(len(exps_centers)), True, np.zeros(len(exps_centers), np.int32))
    return b


def test_ssss_closed_form():
    """(ab|cd) of normalised s primitives against the textbook F0 formula."""
    ec = [(1.3, [0, 0, 0]), (0.7, [0, 0, 1.1]), (0.45, [0.3, -0.8, 0.2]), (2.1, [1.0, 0.5, -0.4])]
    eri = ints.four_index(_s_basis(ec))
    a = [np.array(c) for _, c in ec]
    z = [e for e, _ in ec]
    N = [(2 * x / math.pi) ** 0.75 for x in z]
    p, q = z[0] + z[1], z[2] + z[3]
    P = (z[0] * a[0] + z[1] * a[1]) / p
    Q = (z[2] * a[2] + z[3] * a[3]) / q
    Kab = math.exp(-z[0] * z[1] / p * np.sum((a[0] - a[1]) ** 2))
    Kcd = math.exp(-z[2] * z[3] / q * np.sum((a[2] - a[3]) ** 2))
    T = p * q / (p + q) * np.sum((P - Q) ** 2)
    F0 = 0.5 * math.sqrt(math.pi / T) * erf(math.sqrt(T))
    ref = N[0] * N[1] * N[2] * N[3] * 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * Kab * Kcd * F0
    assert abs(eri[0, 1, 2, 3] - ref) < 1e-14
    # 8-fold permutational symmetry
    assert abs(eri[0, 1, 2, 3] - eri[3, 2, 1, 0]) < 1e-15 and abs(eri[0, 1, 2, 3] - eri[2, 3, 0, 1]) < 1e-15


def test_basis_normalisation_and_sizes():
    mol = water(0.958, 104.4776)
    prim, aux = BasisSet.build(mol, "cc-pvdz"), BasisSet.build(mol, "cc-pvdz-jkfit")
    assert prim.nbf() == 24 and aux.nbf() == 116
    for b in (prim, aux):
        S = ints.overlap(b)
        assert np.abs(np.diag(S) - 1).max() < 1e-13
        assert np.abs(S - S.T).max() == 0.0
        assert np.linalg.eigvalsh(S).min() > 0


def test_kinetic_virial_single_s():
    """<s|T|s> = 3a/2 for a normalised s primitive of exponent a."""
    b = _s_basis([(0.83, [0.1, 0.2, 0.3])])
    assert abs(ints.kinetic(b)[0, 0] - 1.5 * 0.83) < 1e-14


# ------------------------------------------------------------------ metric power
def test_metric_power_identity_and_truncation():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((30, 30))
    J = A @ A.T + 30 * np.eye(30)
    Jm, kept, _ = metric_power(J, -0.5, 1e-10)
    assert kept == 30
    assert np.abs(Jm @ J @ Jm - np.eye(30)).max() < 1e-12
    # constructed spectrum: eigenvalues below condition*max are dropped (matrix.cc:2399-2414)
    U = np.linalg.qr(rng.standard_normal((6, 6)))[0]
    lam = np.array([1.0, 0.5, 0.1, 1e-3, 5e-11, 2e-12])
    Jc = (U * lam) @ U.T
    Jm, kept, _ = metric_power(Jc, -0.5, 1e-10)
    assert kept == 4
    expect = (U[:, :4] * lam[:4] ** -0.5) @ U[:, :4].T
    assert np.abs(Jm - expect).max() < 1e-9


# ------------------------------------------------------------------ J/K against brute force
@pytest.fixture(scope="module")
def h2o2():
    mol = M.water_cluster((2, 1, 1), seed=4)
    prim, aux = BasisSet.build(mol, "cc-pvdz"), BasisSet.build(mol, "cc-pvdz-jkfit")
    o = MemDFJKOracle(prim, aux)
    B = o.dense_B()
    eri = np.einsum("Qmn,Qls->mnls", B, B)
    return mol, prim, aux, o, B, eri


def test_jk_symmetric_vs_fourindex(h2o2):
    mol, prim, aux, o, B, eri = h2o2
    C = np.random.default_rng(1).standard_normal((prim.nbf(), 10)) * 0.2
    D, J, K = o.compute([C])
    Dd = C @ C.T
    assert np.abs(J[0] - np.einsum("mnls,ls->mn", eri, Dd)).max() < 1e-12
    assert np.abs(K[0] - np.einsum("mlns,ls->mn", eri, Dd)).max() < 1e-12
    assert np.abs(K[0] - K[0].T).max() < 1e-13 and np.abs(J[0] - J[0].T).max() < 1e-13
    assert np.vdot(Dd, K[0]) > 0


def test_jk_nonsymmetric_vs_fourindex(h2o2):
    mol, prim, aux, o, B, eri = h2o2
    rng = np.random.default_rng(2)
    L, R = rng.standard_normal((prim.nbf(), 7)), rng.standard_normal((prim.nbf(), 7))
    D, J, K = o.compute([L], [R])
    Dd = L @ R.T
    assert np.abs(J[0] - np.einsum("mnls,ls->mn", eri, Dd)).max() < 1e-11
    assert np.abs(K[0] - np.einsum("mlns,ls->mn", eri, Dd)).max() < 1e-11


def test_linearity_incremental_forms(h2o2):
    """K(D1) - K(D0) = K via [C1,C0]/[C1,-C0] = K via [C1,dC]/[dC,C0] (dC = C1 - C0)."""
    mol, prim, aux, o, B, eri = h2o2
    rng = np.random.default_rng(3)
    C1 = rng.standard_normal((prim.nbf(), 10)) * 0.2
    C0 = C1 + 1e-3 * rng.standard_normal(C1.shape)
    _, J1, K1 = o.compute([C1])
    _, J0, K0 = o.compute([C0])
    _, Js, Ks = o.compute([np.hstack([C1, C0])], [np.hstack([C1, -C0])])
    dC = C1 - C0
    _, Jt, Kt = o.compute([np.hstack([C1, dC])], [np.hstack([dC, C0])])
    for J, K in ((Js, Ks), (Jt, Kt)):
        assert np.abs(K[0] - (K1[0] - K0[0])).max() < 1e-12
        assert np.abs(J[0] - (J1[0] - J0[0])).max() < 1e-12


def test_energy_invariant_under_rotation(h2o2):
    mol, prim, aux, o, B, eri = h2o2
    rng = np.random.default_rng(5)
    C = rng.standard_normal((prim.nbf(), 10)) * 0.2
    U = np.linalg.qr(rng.standard_normal((10, 10)))[0]
    e = []
    for Cx in (C, C @ U):
        D, J, K = o.compute([Cx])
        e.append(np.vdot(D[0], 2 * J[0] - K[0]))
    assert abs(e[0] - e[1]) < 1e-12 * abs(e[0])


def test_adjoint_energy_and_gradients(h2o2):
    """oracle.adjoint: E against the dense four-index contraction and the pinned J/K path;
    dE/dC and dE/dB against central differences of E (E is exactly quadratic in B,
    quartic in C), single elements and symmetric pairs."""
    from oracle.adjoint import energy, energy_adjoint
    mol, prim, aux, o, B, eri = h2o2
    rng = np.random.default_rng(11)
    C = rng.standard_normal((prim.nbf(), 6)) * 0.3
    E, dEdC, Gam = energy_adjoint(B, C)
    Dd = C @ C.T
    E4 = 2 * np.einsum("mn,mnls,ls->", Dd, eri, Dd) - np.einsum("mn,mlns,ls->", Dd, eri, Dd)
    Do, Jo, Ko = o.compute([C])
    assert abs(E - E4) < 1e-11 * abs(E4)
    assert abs(E - np.vdot(Do[0], 2 * Jo[0] - Ko[0])) < 1e-11 * abs(E)
    assert abs(energy(B, C) - E) < 1e-11 * abs(E)
    h = 1e-5
    for _ in range(12):
        m, i = rng.integers(prim.nbf()), rng.integers(6)
        Cp, Cm = C.copy(), C.copy()
        Cp[m, i] += h
        Cm[m, i] -= h
        fd = (energy(B, Cp) - energy(B, Cm)) / (2 * h)
        assert abs(fd - dEdC[m, i]) < 1e-6 * max(1.0, abs(dEdC[m, i]))
    for _ in range(12):
        q, m, n = rng.integers(B.shape[0]), rng.integers(prim.nbf()), rng.integers(prim.nbf())
        Bp, Bm = B.copy(), B.copy()
        Bp[q, m, n] += h
        Bm[q, m, n] -= h
        fd = (energy(Bp, C) - energy(Bm, C)) / (2 * h)
        assert abs(fd - Gam[q, m, n]) < 1e-6 * max(1.0, abs(Gam[q, m, n]))
        if m != n:  # the stored (packed) element moves both halves
            Bp[q, n, m] += h
            Bm[q, n, m] -= h
            fd = (energy(Bp, C) - energy(Bm, C)) / (2 * h)
            assert abs(fd - (Gam[q, m, n] + Gam[q, n, m])) < 1e-6 * max(1.0, abs(Gam[q, m, n]))
    assert np.abs(Gam - Gam.transpose(0, 2, 1)).max() < 1e-12
    _, _, Gs = energy_adjoint(B, C, 5, 17)
    assert np.array_equal(Gs, Gam[5:17])


def test_cholesky_fit_equals_eigen_fit(h2o2):
    """With no eigenvalue truncated, B' = L^-1 (P|mn) gives the same J, K as J^-1/2 (P|mn)."""
    mol, prim, aux, o, B, eri = h2o2
    assert o.n_kept == aux.nbf()
    L = np.linalg.cholesky(o.metric)
    Bc = np.linalg.solve(L, o.Pmn.reshape(aux.nbf(), -1)).reshape(o.Pmn.shape)
    eri_c = np.einsum("Qmn,Qls->mnls", Bc, Bc)
    assert np.abs(eri_c - eri).max() < 1e-11


def test_ppq_layout_roundtrip(h2o2):
    mol, prim, aux, o, B, eri = h2o2
    Bref = np.einsum("QP,Pmn->Qmn", o.metp, o.Pmn)
    Bref[:, ~o.sp.fun_mask] = 0.0
    assert np.abs(ppq_to_dense(o.Ppq, o.sp) - Bref).max() < 1e-13


def test_screening_drops_far_pairs():
    """Two waters 30 bohr apart: cross pairs fall below the Schwarz tolerance (dfhelper.cc:371-387)."""
    mol = M.Molecule.from_string("O 0 0 0\nH 0 0.76 0.59\nH 0 -0.76 0.59\nO 0 0 40\nH 0 0.76 40.59\nH 0 -0.76 40.59")
    prim = BasisSet.build(mol, "cc-pvdz")
    fm, sm, mx = ints.schwarz(prim)
    sp = prepare_sparsity(fm, sm, mx, 1e-12, 10)
    assert not sp.fun_mask[:24, 24:].any()
    assert sp.fun_mask[:24, :24].all() and sp.fun_mask[24:, 24:].all()


# ------------------------------------------------------------------ synthetic generator
def test_philox_known_answers():
    """Random123 philox4x32-10 known-answer vectors."""
    cases = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
             ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
             ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
              (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for c, k, expect in cases:
        out = philox4x32_10(*[np.array([x], np.uint32) for x in c], *k)
        assert tuple(int(o[0]) for o in out) == expect


def test_synthetic_B_shape_and_statistics():
    from oracle.synthetic import synthetic_B
    B = synthetic_B(64, 0, 4, 11)
    assert np.abs(B - B.transpose(0, 2, 1)).max() == 0.0
    # s_Q scaling: diagonal of B_Q / s_Q is N(0,1)-distributed G_Qmm
    g = np.concatenate([np.diag(B[q]) * np.sqrt(q + 1.0) for q in range(4)])
    assert abs(g.mean()) < 0.3 and 0.7 < g.std() < 1.3
