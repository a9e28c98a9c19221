"""DF-RHF and DF-UHF nuclear gradients (psi4_amd/gradient.py) on the CPU, two-electron
part from the oracle adjoint (oracle/adjoint.py) on the oracle's Cholesky-fitted tensor.

Pins:
  * psi4 tests/scf6 (water, R = 1.0 A, A = 104.5, cc-pVDZ, DF with cc-pVDZ-jkfit):
    E(DF-RHF) = -76.02139746388085 (output.ref:277) and the analytic gradient of
    DFJKGrad (output.ref:377-382, 12 digits; psi4 itself compares to 1e-5 and to 5-point
    finite differences at 1e-7);
  * the triplet DF-UHF energy and gradient of the same test (output.ref:6766, 6876-6882);
  * central differences of this build's own SCF energies (5-point, h = 1e-3 bohr);
  * translational and rotational invariance of the total gradient.
The psi4 geometry is its oriented frame (output.ref:40-50, Angstrom), converted with the
CODATA-2006 Bohr radius, which reproduces its E_nuc = 8.801465529972067 to 1e-12.
"""
import numpy as np
import pytest

from helpers import OracleJK

ANG = np.array([[0.0, 0.0, -0.068516219310], [0.0, -0.790689573744, 0.543701060724],
                [0.0, 0.790689573744, 0.543701060724]])
BOHR_2006 = 0.52917720859
PSI4_E = -76.02139746388085
PSI4_GRAD = np.array([[0.0, 0.0, -0.059597382426], [0.0, -0.043039524413, 0.029798691213],
                      [0.0, 0.043039524413, 0.029798691213]])


def _water(xyz_bohr=None):
    import psi4_amd.molecule as M
    return M.Molecule(["O", "H", "H"], ANG / BOHR_2006 if xyz_bohr is None else xyz_bohr)


def _scf(mol):
    from psi4_amd.basisset import BasisSet
    from psi4_amd.scf import RHF
    prim, aux = BasisSet.build(mol, "cc-pvdz"), BasisSet.build(mol, "cc-pvdz-jkfit")
    jk = OracleJK(prim, aux, fit="cholesky")
    wfn = RHF(mol, jk=jk, e_convergence=1e-12, d_convergence=1e-10)
    return wfn, wfn.compute_energy(), jk


def _oracle_adjoint(jk):
    from oracle.adjoint import energy_adjoint
    B = jk.ora.dense_B()

    def adj(Co):
        E, _, G = energy_adjoint(B, Co)
        return E, G, B
    return adj


@pytest.fixture(scope="module")
def scf6():
    from psi4_amd.gradient import rhf_gradient
    mol = _water()
    wfn, res, jk = _scf(mol)
    g, terms = rhf_gradient(wfn, adjoint=_oracle_adjoint(jk))
    return mol, res, g, terms


def test_scf6_energy_and_gradient(scf6):
    mol, res, g, terms = scf6
    assert abs(mol.nuclear_repulsion_energy() - 8.801465529972067) < 1e-11
    assert abs(res.energy - PSI4_E) < 1e-9
    assert np.abs(g - PSI4_GRAD).max() < 1e-8


def test_gradient_invariances(scf6):
    mol, res, g, terms = scf6
    assert np.abs(g.sum(axis=0)).max() < 1e-10                     # translation
    torque = np.cross(mol.xyz, g).sum(axis=0)
    assert np.abs(torque).max() < 1e-10                             # rotation


@pytest.mark.parametrize("atom,axis", [(0, 2), (1, 1), (1, 2)])
def test_gradient_vs_central_differences(scf6, atom, axis):
    mol, res, g, terms = scf6
    h = 1e-3
    e = []
    for k in (-2, -1, 1, 2):
        xyz = mol.xyz.copy()
        xyz[atom, axis] += k * h
        e.append(_scf(_water(xyz))[1].energy)
    fd = (e[0] - 8 * e[1] + 8 * e[2] - e[3]) / (12 * h)
    assert abs(fd - g[atom, axis]) < 1e-7, (fd, g[atom, axis])


PSI4_E_UHF = -75.79200445921231          # scf6 output.ref:6766 (triplet, DF-UHF)
PSI4_GRAD_UHF = np.array([[0.0, 0.0, 0.070328030317], [0.0, 0.046340283191, -0.035164015159],
                          [0.0, -0.046340283191, -0.035164015159]])   # output.ref:6876-6882


@pytest.fixture(scope="module")
def scf6_uhf():
    """Triplet water of scf6 (Nalpha 6, Nbeta 4); UHF started from the RHF orbitals of the
    singlet (psi4 starts from SAD; the core guess lands on an excited UHF solution)."""
    import psi4_amd.molecule as M
    from psi4_amd.basisset import BasisSet
    from psi4_amd.gradient import uhf_gradient
    from psi4_amd.scf import RHF
    from psi4_amd.uhf import UHF
    xyz = ANG / BOHR_2006
    mol = M.Molecule(["O", "H", "H"], xyz, multiplicity=3)
    prim, aux = BasisSet.build(mol, "cc-pvdz"), BasisSet.build(mol, "cc-pvdz-jkfit")
    jk = OracleJK(prim, aux, fit="cholesky")
    guess = RHF(M.Molecule(["O", "H", "H"], xyz), jk=jk).compute_energy().C
    wfn = UHF(mol, jk=jk, e_convergence=1e-12, d_convergence=1e-10, guess=guess)
    res = wfn.compute_energy()
    g, terms = uhf_gradient(wfn, adjoint=_oracle_adjoint(jk))
    return mol, wfn, res, g, jk, guess


def test_scf6_uhf_energy_and_gradient(scf6_uhf):
    mol, wfn, res, g, jk, guess = scf6_uhf
    assert (wfn.nalpha, wfn.nbeta) == (6, 4)
    assert abs(res.energy - PSI4_E_UHF) < 1e-9
    assert np.abs(g - PSI4_GRAD_UHF).max() < 2e-8
    assert np.abs(g.sum(axis=0)).max() < 1e-10


def test_uhf_gradient_vs_central_differences(scf6_uhf):
    import psi4_amd.molecule as M
    from psi4_amd.basisset import BasisSet
    from psi4_amd.uhf import UHF
    mol, wfn, res, g, jk, guess = scf6_uhf
    h = 1e-3
    e = []
    for k in (-2, -1, 1, 2):
        xyz = mol.xyz.copy()
        xyz[1, 2] += k * h
        m = M.Molecule(["O", "H", "H"], xyz, multiplicity=3)
        j = OracleJK(BasisSet.build(m, "cc-pvdz"), BasisSet.build(m, "cc-pvdz-jkfit"), fit="cholesky")
        e.append(UHF(m, jk=j, e_convergence=1e-12, d_convergence=1e-10, guess=guess).compute_energy().energy)
    fd = (e[0] - 8 * e[1] + 8 * e[2] - e[3]) / (12 * h)
    assert abs(fd - g[1, 2]) < 1e-7, (fd, g[1, 2])
