"""psi4's eigen J^-1/2 fit with the condition cut (DFJK_FIT_EIGEN) on the device.

Reference: libmints/matrix.cc:2370-2424 (Matrix::power: eigenvalues |a| < condition * max|a|
and non-finite powers give 0), applied by lib3index/dfhelper.cc:1491-1517.  The oracle is
oracle/dfhelper.py's metric_power and MemDFJKOracle(fit="eigen", condition=...).

* A truncating condition on (H2O)_4: the number of eigenvalues kept, the fitted tensor and
  J/K (symmetric and non-symmetric pairs) match the oracle's eigen fit with the same
  condition (B to 1e-10, J/K to 1e-11 of their largest element).  The condition is placed between
  two eigenvalues that are well separated, so the two eigensolvers agree on the cut.
* Without truncation the eigen fit and the Cholesky fit differ by an orthogonal rotation of
  the auxiliary index, which J and K do not see: J/K of the two device fits agree to 1e-11.
"""
import numpy as np
import pytest

from conftest import announce
from helpers import maxabs, random_orbitals, water_system

pytestmark = pytest.mark.gpu


def _jk(prim, aux, fit, condition=None):
    from psi4_amd.core import FIT_CHOLESKY, FIT_EIGEN, JK
    jk = JK.build(prim, aux)
    jk.set_print(0)
    jk.set_fitting(FIT_EIGEN if fit == "eigen" else FIT_CHOLESKY)
    if condition is not None:
        jk.set_condition(condition)
    jk.initialize()
    return jk


def _separated_condition(eigs, target=1e-5):
    """A condition between two consecutive metric eigenvalue ratios near target whose ratio
    is at least 1.5 (no eigenvalue within rounding of the cut)."""
    r = np.sort(np.abs(eigs)) / np.max(np.abs(eigs))
    best = None
    for i in range(len(r) - 1):
        if r[i + 1] / max(r[i], 1e-300) >= 1.5:
            c = np.sqrt(r[i] * r[i + 1])
            if best is None or abs(np.log(c / target)) < abs(np.log(best / target)):
                best = c
    return float(best)


def test_eigen_fit_truncating_matches_oracle(gpu_available):
    from oracle.dfhelper import MemDFJKOracle, metric_power
    from psi4_amd import ints
    mol, prim, aux = water_system((2, 2, 1), seed=0)
    Jm = ints.metric(aux)
    eigs = np.linalg.eigvalsh(Jm)
    cond = _separated_condition(eigs)
    _, n_kept, _ = metric_power(Jm, -0.5, cond)
    assert n_kept < aux.nbf()                     # the cut does bite
    announce(f"(H2O)_4 eigen fit, condition {cond:.3e} keeps {n_kept} of {aux.nbf()}")
    jk = _jk(prim, aux, "eigen", cond)
    info = jk.info()
    assert info["fitting"] == 1 and info["naux_kept"] == n_kept
    ora = MemDFJKOracle(prim, aux, fit="eigen", condition=cond)
    B = jk.get_B_dense(0, aux.nbf())
    Bo = ora.dense_B()
    eb = maxabs(B, Bo) / np.abs(Bo).max()
    C = random_orbitals(prim.nbf(), 20, 11)
    Cr = random_orbitals(prim.nbf(), 20, 12)
    jk.C_clear()
    jk.C_left_add(C)
    jk.compute()
    _, Jo, Ko = ora.compute([C])
    ej, ek = maxabs(jk.J()[0], Jo[0]), maxabs(jk.K()[0], Ko[0])
    jk.C_clear()
    jk.C_left_add(C)
    jk.C_right_add(Cr)
    jk.compute()
    _, Jn, Kn = ora.compute([C], [Cr])
    ejn, ekn = maxabs(jk.J()[0], Jn[0]), maxabs(jk.K()[0], Kn[0])
    print(f"  B rel {eb:.2e}  J {ej:.2e} K {ek:.2e}  non-symmetric J {ejn:.2e} K {ekn:.2e}", flush=True)
    assert eb < 1e-10
    assert max(ej, ejn) < 1e-11 * max(1.0, np.abs(Jo[0]).max())
    assert max(ek, ekn) < 1e-11 * max(1.0, np.abs(Ko[0]).max())


def test_eigen_fit_equals_cholesky_fit_without_truncation(gpu_available):
    mol, prim, aux = water_system((2, 2, 1), seed=0)
    announce("(H2O)_4 eigen vs Cholesky fit, default condition (nothing cut)")
    je = _jk(prim, aux, "eigen")
    jc = _jk(prim, aux, "cholesky")
    assert je.info()["naux_kept"] == aux.nbf()
    C = random_orbitals(prim.nbf(), 20, 5)
    out = []
    for jk in (je, jc):
        jk.C_clear()
        jk.C_left_add(C)
        jk.compute()
        out.append((jk.J()[0].copy(), jk.K()[0].copy()))
    ej, ek = maxabs(out[0][0], out[1][0]), maxabs(out[0][1], out[1][1])
    print(f"  eigen vs Cholesky fit: J {ej:.2e} K {ek:.2e}", flush=True)
    assert ej < 1e-11 and ek < 1e-11
