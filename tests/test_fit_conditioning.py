"""How reproducible psi4's FP64 fit of the Coulomb metric is (CPU, no GPU).

psi4 forms J^-1/2 with LAPACK DSYEV (libmints/matrix.cc:2370-2424, applied by
lib3index/dfhelper.cc:1491-1517).  On a water cluster the metric's condition number is
about 1e7, so any two FP64 factorisations of the same metric -- DSYEV, DSYEVD, DSYEVR
(all nothing truncated) or a Cholesky factor -- give two-electron energies that differ by
a few 1e-10 Eh at the same orbitals.  That is the floor of any energy comparison against
"the" eigen fit: the device engine's contractions are checked on its own tensor at 1e-10 Eh
(and J/K at 1e-11), and its fit against the oracle's at the 1e-9 Eh this spread allows.

E_2e = 2 |W c|^2 - |W T|^2 for J^-1 ~ W^T W, c_P = (P|mn) D_mn, T_P,ij = (P|ij) with
(P|ij) = C^T (P|mn) C: exactly the energy the fitted tensor B = W (P|mn) gives.
"""
import numpy as np
import scipy.linalg as sl

from helpers import water_system


def _energies(grid):
    from psi4_amd import ints
    mol, prim, aux = water_system(grid, seed=0)
    nbf, naux, nocc = prim.nbf(), aux.nbf(), mol.nelectron() // 2
    Jm = ints.metric(aux)
    P = ints.three_index(prim, aux)
    S = ints.overlap(prim)
    s, U = np.linalg.eigh(S)
    C = ((U / np.sqrt(s)) @ U.T) @ np.linalg.qr(np.random.default_rng(1).standard_normal((nbf, nocc)))[0]
    c = np.einsum("pmn,mn->p", P, C @ C.T)
    T = np.einsum("pmn,mi,nj->pij", P, C, C, optimize=True).reshape(naux, -1)
    R = np.hstack([c[:, None], T])
    del P

    def e2(W):
        y = W @ R
        return 2.0 * float(y[:, 0] @ y[:, 0]) - float(np.sum(y[:, 1:] ** 2))

    out = {}
    for driver in ("ev", "evd", "evr"):
        a, V = sl.eigh(Jm, driver=driver)
        out["dsy" + driver] = e2((V / np.sqrt(a)) @ V.T)
    L = np.linalg.cholesky(Jm)
    out["cholesky"] = e2(sl.solve_triangular(L, np.eye(naux), lower=True))
    return out


def test_fp64_metric_fits_agree_only_to_a_few_1e10():
    """(H2O)_16 cc-pVDZ/cc-pVDZ-jkfit (nbf 384, naux 1856): every fit within 1e-9 Eh of psi4's
    DSYEV fit, while two LAPACK eigen drivers already disagree by more than 1e-10 Eh."""
    E = _energies((2, 4, 2))
    v = np.array(list(E.values()))
    spread = float(v.max() - v.min())
    print("  E_2e by fit:", {k: f"{x:.12f}" for k, x in E.items()}, f"spread {spread:.2e}")
    assert all(abs(x - E["dsyev"]) < 1e-9 for x in v)
    assert spread > 1e-10
