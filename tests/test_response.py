"""CPHF driver (psi4_amd/response.py) on the CPU, with the oracle JK behind it.

Pins, for water / aug-cc-pVDZ / aug-cc-pVDZ-jkfit at psi4's geometry of
tests/scf-response1 (output.ref:131-135):
  * DF-RHF energy -76.00326530575171 (output.ref:267; psi4 stopped at a density
    RMS of 7e-7, so agreement is to ~1e-9) and E_nuc 8.001332299975491;
  * the dipole moment 0.8260 e a0 (output.ref:288-289) -- pins the dipole integrals;
  * the static dipole polarizability, psi4's bench values
    (input.dat:14-19, compared to 3 decimals as psi4 does) through the principal
    values (frame independent);
  * the same tensor from finite dipole fields (second differences of the SCF
    energy, Richardson-extrapolated), a check that involves no CPHF code at all.
"""
import numpy as np
import pytest

from helpers import OracleJK

BOHR_2006 = 0.52917720859
XYZ_A = np.array([[0.0, 0.0, -0.075367841241], [0.0, -0.869758531118, 0.598071166797],
                  [0.0, 0.869758531118, 0.598071166797]])
PSI4_E = -76.00326530575171
PSI4_ALPHA = np.array([8.01603, 12.52796, 10.02099])   # XX, YY, ZZ; off-diagonals 0 to 5 decimals


def _water():
    import psi4_amd.molecule as M
    return M.Molecule(["O", "H", "H"], XYZ_A / BOHR_2006)


@pytest.fixture(scope="module")
def water_aug():
    from psi4_amd.basisset import BasisSet
    from psi4_amd.scf import RHF
    mol = _water()
    prim = BasisSet.build(mol, "aug-cc-pvdz")
    aux = BasisSet.build(mol, "aug-cc-pvdz-jkfit")
    jk = OracleJK(prim, aux)
    wfn = RHF(mol, basis="aug-cc-pvdz", jk=jk, e_convergence=1e-11, d_convergence=1e-9)
    res = wfn.compute_energy()
    return mol, wfn, res, jk


def test_scf_energy_and_dipole(water_aug):
    from psi4_amd import ints
    mol, wfn, res, _ = water_aug
    assert wfn.basis.nbf() == 41 and wfn.aux.nbf() == 150
    assert abs(mol.nuclear_repulsion_energy() - 8.001332299975491) < 1e-12
    assert abs(res.energy - PSI4_E) < 2e-9
    mu = ints.dipole(wfn.basis)
    dip = -2.0 * np.einsum("dmn,mn->d", mu, res.D) + (mol.Z[:, None] * mol.xyz).sum(0)
    assert abs(np.linalg.norm(dip) - 0.8260) < 6e-5
    assert abs(dip[0]) < 1e-10 and abs(dip[1]) < 1e-10


def test_dipole_polarizability_vs_psi4(water_aug):
    from psi4_amd.response import cpscf_linear_response
    _, wfn, _, _ = water_aug
    (alpha,) = cpscf_linear_response(wfn, "DIPOLE_POLARIZABILITIES", conv_tol=1e-8, max_iter=50)
    assert wfn.cphf_converged
    assert np.abs(alpha - alpha.T).max() < 1e-6
    w = np.sort(np.linalg.eigvalsh(0.5 * (alpha + alpha.T)))
    assert np.abs(w - np.sort(PSI4_ALPHA)).max() < 1e-3
    # psi4's frame equals the input frame here (C2 axis z, molecule in the yz plane)
    assert np.abs(np.diag(alpha) - PSI4_ALPHA).max() < 1e-3


def test_cphf_residual_and_default_settings(water_aug):
    """psi4's defaults (conv_tol 1e-5, max_iter 10) still land within psi4's 1e-3 bar, and the
    converged solution satisfies the Hessian equation it was solved for."""
    from psi4_amd import ints
    from psi4_amd.response import cpscf_linear_response
    _, wfn, _, _ = water_aug
    (alpha,) = cpscf_linear_response(wfn, "DIPOLE_POLARIZABILITIES")
    assert np.abs(np.diag(alpha) - PSI4_ALPHA).max() < 1e-3
    Co, Cv = wfn.Ca[:, :wfn.nocc], wfn.Ca[:, wfn.nocc:]
    b = [Co.T @ m @ Cv for m in ints.dipole(wfn.basis)]
    x = wfn.cphf_solve(b, 1e-9, 60)
    Hx = wfn.cphf_Hx(x)
    for bi, hi in zip(b, Hx):
        assert np.sqrt(np.sum((bi - hi) ** 2) / np.sum(bi * bi)) < 1e-9
    with pytest.raises(ValueError, match="occ x vir"):
        wfn.onel_Hx([np.zeros((3, 3))])


def test_polarizability_vs_finite_field(water_aug):
    """alpha from CPHF against central second differences of the SCF energy in a dipole
    field (psi4's PERTURB_DIPOLE), alpha_dd = -(E(+l) + E(-l) - 2E(0)) / l^2 extrapolated to l = 0."""
    from psi4_amd.response import cpscf_linear_response
    from psi4_amd.scf import RHF
    mol, wfn, res, jk = water_aug
    (alpha,) = cpscf_linear_response(wfn, "DIPOLE_POLARIZABILITIES", conv_tol=1e-9, max_iter=60)
    lam = 2e-3

    def energy_in_field(v):
        return RHF(mol, basis="aug-cc-pvdz", jk=jk, e_convergence=1e-12, d_convergence=1e-10,
                   perturb_dipole=v).compute_energy().energy

    e0 = res.energy

    def second_difference(d, step):
        v = np.zeros(3)
        v[d] = step
        return -(energy_in_field(v) + energy_in_field(-v) - 2 * e0) / step ** 2

    for d in range(3):
        # Richardson: the O(lambda^2) hyperpolarizability term cancels
        a_ff = (4.0 * second_difference(d, lam) - second_difference(d, 2 * lam)) / 3.0
        print(f"  alpha[{d}{d}] CPHF {alpha[d, d]:.7f}  finite field {a_ff:.7f}", flush=True)
        assert abs(a_ff - alpha[d, d]) < 2e-5, (d, a_ff, alpha[d, d])
