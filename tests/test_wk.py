"""Range-separated exchange (wK) on the CPU: the erf-attenuated integrals of the host
engine and the oracle's restatement of psi4's non-combined DF wK
(lib3index/dfhelper.cc:589-699, 3378-3438).

Integral pins: the closed form of the attenuated Coulomb interaction of two s-type
Gaussian charge distributions, erf(sqrt(a') R)/R with 1/a' = 1/p + 1/q + 1/omega^2
(off-centre and same-centre), and the omega -> infinity limit (plain Coulomb) of the
three- and four-index sets.  Oracle pins: omega -> infinity turns wK into K; the
pQq-layout restatement equals the dense sum_PQ ((P|mi) C_l) [J^-1]_PQ ((Q|w|ni) C_r)
with C_l != C_r (catches a swapped side, a transposed product or a layout slip); and
DF wK approaches the exact four-index erf exchange within the fitting error.
"""
import numpy as np
import pytest
from scipy.special import erf

import psi4_amd.molecule as M
from oracle.dfhelper import MemDFJKOracle
from psi4_amd import ints
from psi4_amd.basisset import BasisSet

OMEGA = 0.4


def _sbasis(exps, centers):
    """One single-primitive s shell per exponent (unit coefficient, normalised by the engine)."""
    n = len(exps)
    return BasisSet("s-test", np.zeros(n, np.int32), np.ones(n, np.int32), np.asarray(centers, float).reshape(-1, 3),
                    np.asarray(exps, float), np.ones(n), True, np.zeros(n, np.int32))


def _charge(c):
    """Integral of a unit-normalised s Gaussian of exponent c."""
    return (2.0 * c / np.pi) ** 0.75 * (np.pi / c) ** 1.5


@pytest.mark.parametrize("R", [0.0, 1.37])
def test_erf_three_index_closed_form_s(R):
    a, b = 0.8, 1.3
    A, B = np.zeros(3), np.array([0.3, -0.4, 1.1]) * (R / np.linalg.norm([0.3, -0.4, 1.1]) if R else 0.0)
    aux, prim = _sbasis([a], [A]), _sbasis([b], [B])
    # the pair m*m is a Gaussian of exponent 2b with charge N_b^2 (pi/2b)^1.5
    qa = _charge(a)
    qmm = (2.0 * b / np.pi) ** 1.5 * (np.pi / (2.0 * b)) ** 1.5
    for omega in (None, OMEGA, 2.5):
        inv = 1.0 / a + 1.0 / (2.0 * b) + (1.0 / omega ** 2 if omega else 0.0)
        ap = 1.0 / inv
        exact = qa * qmm * (2.0 * np.sqrt(ap / np.pi) if R == 0.0 else erf(np.sqrt(ap) * R) / R)
        got = ints.three_index(prim, aux, omega=omega)[0, 0, 0]
        assert abs(got - exact) < 1e-13 * abs(exact), (omega, got, exact)


def test_erf_integrals_large_omega_limit():
    mol = M.water_zmat(0.96, 104.5)
    prim = BasisSet.build(mol, "cc-pvdz")
    aux = BasisSet.build(mol, "cc-pvdz-jkfit")
    # a' = a w^2/(a + w^2): relative change ~ a/w^2 (a up to ~1e4 for the tight O exponents)
    P = ints.three_index(prim, aux)
    Pw = ints.three_index(prim, aux, omega=1e8)
    assert np.abs(Pw - P).max() < 1e-9 * np.abs(P).max()
    Jm, Jw = ints.metric(aux), ints.metric(aux, omega=1e8)
    assert np.abs(Jw - Jm).max() < 1e-9 * np.abs(Jm).max()
    E, Ew = ints.four_index(prim), ints.four_index(prim, omega=1e8)
    assert np.abs(Ew - E).max() < 1e-9 * np.abs(E).max()
    # a finite omega attenuates: (mn|w|mn) < (mn|mn) on the diagonal, symmetric like 1/r
    E4 = ints.four_index(prim, omega=OMEGA)
    n = prim.nbf()
    d, dw = np.einsum("mnmn->mn", E), np.einsum("mnmn->mn", E4)
    assert np.all(dw <= d + 1e-14) and np.all(dw > 0)
    assert np.abs(E4 - E4.transpose(2, 3, 0, 1)).max() < 1e-13
    assert np.abs(E4 - E4.transpose(1, 0, 2, 3)).max() < 1e-13
    assert n == 24


def _water_oracle(omega, fit="eigen"):
    mol = M.water_zmat(0.96, 104.5)
    prim = BasisSet.build(mol, "cc-pvdz")
    aux = BasisSet.build(mol, "cc-pvdz-jkfit")
    return prim, aux, MemDFJKOracle(prim, aux, fit=fit, omega=omega)


def test_oracle_wk_large_omega_is_K():
    prim, aux, o = _water_oracle(1e8)
    rng = np.random.default_rng(5)
    C = np.linalg.qr(rng.standard_normal((prim.nbf(), prim.nbf())))[0][:, :5]
    wK = o.compute_wK([C])[0]
    K = o.compute([C])[2][0]
    assert np.abs(wK - K).max() < 1e-8 * np.abs(K).max()


@pytest.mark.parametrize("fit", ["eigen", "cholesky"])
def test_oracle_wk_equals_dense_route(fit):
    prim, aux, o = _water_oracle(OMEGA, fit)
    rng = np.random.default_rng(6)
    n = prim.nbf()
    Cl, Cr = rng.standard_normal((n, 4)), rng.standard_normal((n, 4))
    wK = o.compute_wK([Cl], [Cr])[0]
    P = ints.three_index(prim, aux)
    W = ints.three_index(prim, aux, omega=OMEGA)
    mask = o.sp.fun_mask
    P[:, ~mask] = 0.0
    W[:, ~mask] = 0.0
    Jinv = np.linalg.inv(ints.metric(aux))
    L = np.einsum("pml,li->pmi", P, Cl)
    R = np.einsum("qns,si->qni", W, Cr)
    dense = np.einsum("pmi,pq,qni->mn", L, Jinv, R)
    assert np.abs(wK - dense).max() < 1e-10 * np.abs(dense).max()
    # not symmetric in general, and left/right are not interchangeable
    assert np.abs(wK - wK.T).max() > 1e-6 * np.abs(wK).max()
    swapped = o.compute_wK([Cr], [Cl])[0]
    assert np.abs(swapped - wK).max() > 1e-6 * np.abs(wK).max()


def test_oracle_wk_near_exact_four_index():
    prim, aux, o = _water_oracle(OMEGA)
    mol = prim.molecule
    from oracle.rhf import rhf
    S = ints.overlap(prim)
    H = ints.kinetic(prim) + ints.potential(prim, mol)

    def jk(C):
        _, J, K = o.compute([C])
        return J[0], K[0]
    C = rhf(H, S, mol.nelectron() // 2, mol.nuclear_repulsion_energy(), jk)["C"][:, :mol.nelectron() // 2]
    wK = o.compute_wK([C])[0]
    E4 = ints.four_index(prim, omega=OMEGA)
    exact = np.einsum("mlns,ls->mn", E4, C @ C.T)
    # density fitting error of the jkfit basis, a few 1e-4 of the largest element
    assert np.abs(wK - exact).max() < 2e-3 * np.abs(exact).max()
    # and far from the Coulomb exchange: the attenuation is really applied
    K = o.compute([C])[2][0]
    assert np.abs(K - exact).max() > 0.1 * np.abs(exact).max()


def test_oracle_wk_symmetric_pair_is_hermitivitized():
    """lr_symmetric (no C_right): MemDFJK::compute_JK hermitivitizes wK (MemDFJK.cc:101-107);
    with C_right given explicitly (= C_left) the same product is returned as built, and that
    one is not symmetric, so the switch is visible."""
    prim, aux, o = _water_oracle(OMEGA, "cholesky")
    rng = np.random.default_rng(8)
    C = np.linalg.qr(rng.standard_normal((prim.nbf(), prim.nbf())))[0][:, :5]
    sym = o.compute_wK([C])[0]
    raw = o.compute_wK([C], [C])[0]
    assert np.abs(raw - raw.T).max() > 1e-6 * np.abs(raw).max()
    assert np.array_equal(sym, sym.T)
    assert np.abs(sym - 0.5 * (raw + raw.T)).max() < 1e-15 * np.abs(raw).max()


def test_oracle_combined_exchange():
    """wcombine (dfhelper.cc:659-667, 1411, 3022): K is not built and wK pairs J^-1 (Q|mn) C_l
    with (alpha (Q|mn) + beta (Q|w|mn)) C_r.  (alpha, beta) = (1, 0) must give the plain K of
    the J^-1/2 route (different algebra: the raw integrals against the full inverse metric),
    (0, 1) the non-combined wK, and a general pair their combination."""
    mol = M.water_zmat(0.96, 104.5)
    prim = BasisSet.build(mol, "cc-pvdz")
    aux = BasisSet.build(mol, "cc-pvdz-jkfit")
    rng = np.random.default_rng(9)
    n = prim.nbf()
    Cl, Cr = rng.standard_normal((n, 4)), rng.standard_normal((n, 4))
    plain = MemDFJKOracle(prim, aux, omega=OMEGA)
    K = plain.compute([Cl], [Cr])[2][0]
    wK = plain.compute_wK([Cl], [Cr])[0]
    c10 = MemDFJKOracle(prim, aux, omega=OMEGA, wcombine=True, omega_alpha=1.0, omega_beta=0.0)
    assert np.abs(c10.compute_wK([Cl], [Cr])[0] - K).max() < 1e-10 * np.abs(K).max()
    assert not c10.compute([Cl], [Cr])[2][0].any()                 # K not built
    J = plain.compute([Cl], [Cr])[1][0]
    assert np.abs(c10.compute([Cl], [Cr])[1][0] - J).max() < 1e-13 * np.abs(J).max()
    c01 = MemDFJKOracle(prim, aux, omega=OMEGA, wcombine=True, omega_alpha=0.0, omega_beta=1.0)
    assert np.abs(c01.compute_wK([Cl], [Cr])[0] - wK).max() < 1e-13 * np.abs(wK).max()
    a, b = 0.75, -0.4
    cab = MemDFJKOracle(prim, aux, omega=OMEGA, wcombine=True, omega_alpha=a, omega_beta=b)
    ref = a * K + b * wK
    assert np.abs(cab.compute_wK([Cl], [Cr])[0] - ref).max() < 1e-10 * np.abs(ref).max()
