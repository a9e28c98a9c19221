"""CPHF on the GPU JK (multi-pair non-symmetric path inside RHF::twoel_Hx_full).

* water / aug-cc-pVDZ (psi4 tests/scf-response1): polarizability within psi4's
  own 1e-3 bar of its stored values (input.dat:14-19);
* the GPU 4J - K - K^T products against the oracle JK on the device's own B
  (1e-11 of the largest element), and the polarizability of (H2O)_8 / cc-pVDZ
  against the same CPHF driver run with the oracle JK (1e-8).
"""
import numpy as np
import pytest

from conftest import announce
from helpers import OracleJK

pytestmark = pytest.mark.gpu

PSI4_ALPHA = np.array([8.01603, 12.52796, 10.02099])


def _rhf(mol, basis, jk=None):
    from psi4_amd.scf import RHF
    wfn = RHF(mol, basis=basis, jk=jk, e_convergence=1e-11, d_convergence=1e-9)
    wfn.compute_energy()
    return wfn


def test_water_aug_polarizability_gpu(gpu_available):
    import psi4_amd.molecule as M
    from psi4_amd.response import cpscf_linear_response
    xyz = np.array([[0.0, 0.0, -0.075367841241], [0.0, -0.869758531118, 0.598071166797],
                    [0.0, 0.869758531118, 0.598071166797]])
    mol = M.Molecule(["O", "H", "H"], xyz / 0.52917720859)
    announce("water aug-cc-pVDZ RHF + CPHF polarizability on the GPU JK")
    wfn = _rhf(mol, "aug-cc-pvdz")
    (alpha,) = cpscf_linear_response(wfn, "DIPOLE_POLARIZABILITIES", conv_tol=1e-8, max_iter=50)
    print(f"  alpha diag {np.diag(alpha)}", flush=True)
    assert wfn.cphf_converged
    assert np.abs(np.diag(alpha) - PSI4_ALPHA).max() < 1e-3


def test_h2o8_twoel_and_polarizability_vs_oracle(gpu_available):
    from psi4_amd.molecule import water_cluster
    from psi4_amd.response import cpscf_linear_response
    mol = water_cluster((2, 2, 2), seed=3)
    announce("(H2O)_8 cc-pVDZ RHF on the GPU JK")
    wfn = _rhf(mol, "cc-pvdz")
    jk = wfn.jk()
    B = jk.get_B_dense(0, jk.info()["naux"])
    ora = OracleJK(wfn.basis, wfn.aux, B=B)
    del B

    class OracleSide:
        """The same converged reference, with the oracle JK behind the products."""
        def __init__(self, w):
            self.Ca, self.Fa, self.nocc, self.basis = w.Ca, w.Fa, w.nocc, w.basis

        def jk(self):
            return ora

    ref = OracleSide(wfn)
    nocc, nvir = wfn.nocc, wfn.Ca.shape[1] - wfn.nocc
    rng = np.random.default_rng(8)
    xs = [rng.standard_normal((nocc, nvir)) * 0.1 for _ in range(4)]
    from psi4_amd import response
    announce("(H2O)_8 twoel_Hx_full, 4 vectors in one JK call, GPU vs oracle")
    for combine in (True, False):
        got = wfn.twoel_Hx_full(xs, combine, "SO")
        exp = response.twoel_Hx_full(ref, xs, combine, "SO")
        for g, e in zip(got, exp):
            err = np.abs(g - e).max() / max(1.0, np.abs(e).max())
            assert err < 1e-11, (combine, err)
    announce("(H2O)_8 CPHF polarizability, GPU vs oracle JK")
    (a_gpu,) = cpscf_linear_response(wfn, "DIPOLE_POLARIZABILITIES", conv_tol=1e-9, max_iter=60)
    (a_ora,) = response.cpscf_linear_response(ref, "DIPOLE_POLARIZABILITIES", conv_tol=1e-9, max_iter=60)
    print(f"  alpha GPU {np.diag(a_gpu)}  oracle {np.diag(a_ora)}", flush=True)
    assert np.abs(a_gpu - a_ora).max() < 1e-8 * np.abs(a_ora).max()
