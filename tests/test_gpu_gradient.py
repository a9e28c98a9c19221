"""DF-RHF gradient with the two-electron part from the GPU adjoint (dfjk_energy_adjoint).

* psi4 tests/scf6 water: gradient against psi4's analytic DF gradient
  (output.ref:377-382) at 1e-8, energy at 1e-9;
* (H2O)_4 cc-pVDZ: the gradient from the device Gamma equals the one from the oracle
  adjoint on the device's own fitted tensor (same host derivative integrals) to 1e-9,
  and the total force and torque vanish.
"""
import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

ANG = np.array([[0.0, 0.0, -0.068516219310], [0.0, -0.790689573744, 0.543701060724],
                [0.0, 0.790689573744, 0.543701060724]])
PSI4_GRAD = np.array([[0.0, 0.0, -0.059597382426], [0.0, -0.043039524413, 0.029798691213],
                      [0.0, 0.043039524413, 0.029798691213]])


def test_scf6_gradient_on_gpu(gpu_available):
    import psi4_amd.molecule as M
    from psi4_amd.gradient import rhf_gradient
    from psi4_amd.scf import RHF
    mol = M.Molecule(["O", "H", "H"], ANG / 0.52917720859)
    announce("scf6 water DF-RHF gradient, two-electron part from the GPU adjoint")
    wfn = RHF(mol, e_convergence=1e-12, d_convergence=1e-10)
    res = wfn.compute_energy()
    g, _ = rhf_gradient(wfn)
    print(f"  E {res.energy:.12f}  max |g - psi4| {np.abs(g - PSI4_GRAD).max():.2e}", flush=True)
    assert abs(res.energy - (-76.02139746388085)) < 1e-9
    assert np.abs(g - PSI4_GRAD).max() < 1e-8


def test_h2o4_gradient_device_vs_oracle_adjoint(gpu_available):
    from oracle.adjoint import energy_adjoint
    from psi4_amd.gradient import rhf_gradient
    from psi4_amd.molecule import water_cluster
    from psi4_amd.scf import RHF
    mol = water_cluster((2, 2, 1), seed=0)
    announce("(H2O)_4 DF-RHF gradient, device vs oracle adjoint")
    wfn = RHF(mol, e_convergence=1e-11, d_convergence=1e-9)
    wfn.compute_energy()
    g, terms = rhf_gradient(wfn)
    jk = wfn.jk()
    B = jk.get_B_dense(0, jk.info()["naux"])

    def adj(Co):
        E, _, G = energy_adjoint(B, Co)
        return E, G, B
    go, to = rhf_gradient(wfn, adjoint=adj)
    err = np.abs(g - go).max()
    print(f"  max |g_device - g_oracle| {err:.2e}  |g| {np.abs(g).max():.3e}", flush=True)
    assert err < 1e-9
    assert abs(terms["E_2"] - to["E_2"]) < 1e-11 * abs(to["E_2"])
    assert np.abs(g.sum(axis=0)).max() < 1e-9
    # torque-free up to the SCF convergence (d_convergence 1e-9) and the 1e-12 Schwarz screening
    assert np.abs(np.cross(mol.xyz, g).sum(axis=0)).max() < 1e-7


def test_scf6_uhf_energy_and_gradient_on_gpu(gpu_available):
    """Triplet DF-UHF (two density pairs per JK call) and its gradient (two device adjoints)
    against psi4 scf6 (output.ref:6766, 6876-6882)."""
    import psi4_amd.molecule as M
    from psi4_amd.gradient import uhf_gradient
    from psi4_amd.scf import RHF
    from psi4_amd.uhf import UHF
    xyz = ANG / 0.52917720859
    announce("scf6 triplet water DF-UHF energy and gradient on the GPU JK")
    guess = RHF(M.Molecule(["O", "H", "H"], xyz)).compute_energy().C
    wfn = UHF(M.Molecule(["O", "H", "H"], xyz, multiplicity=3), e_convergence=1e-12, d_convergence=1e-10,
              guess=guess)
    res = wfn.compute_energy()
    g, _ = uhf_gradient(wfn)
    ref = np.array([[0.0, 0.0, 0.070328030317], [0.0, 0.046340283191, -0.035164015159],
                    [0.0, -0.046340283191, -0.035164015159]])
    print(f"  E {res.energy:.12f}  max |g - psi4| {np.abs(g - ref).max():.2e}", flush=True)
    assert abs(res.energy - (-75.79200445921231)) < 1e-9
    assert np.abs(g - ref).max() < 2e-8
