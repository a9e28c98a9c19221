"""C5: full RHF on a seeded (H2O)_48 cluster (4x4x3 grid), cc-pVDZ/cc-pVDZ-jkfit
(nbf 1152, naux 5568, nocc 240) on the device JK.

Reference path: libscf_solver/rhf.cc:187-367 and scf_iterator.py:264-609, with the
incremental-Fock pattern of libfock/DirectJK.cc:333-349.

* plain FP64 RHF: the converged energy equals the energy the oracle's contractions
  (oracle/adjoint.py energy, pinned to the dfhelper restatement) give on the device's own
  fitted tensor at the converged orbitals, summed over Q-slices of 256 (J and K, and so
  E = tr(D(2J - K)), are additive over Q): |dE| < 1e-8 Eh;
* incremental BF16x3 RHF with FP64 finishing iterations on the same JK: |E_mixed - E_plain|
  < 1e-8 Eh, and every mixed build within 1e-6 Eh of the FP64 build of the same orbitals;
* the device adjoint at the converged orbitals: <C, dE/dC> = 4E and, summed over Q-slices,
  <B, dE/dB> = 2E (E is homogeneous of degree 4 in C and 2 in B), to 1e-12 relative.
"""
import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

SLICE = 256


@pytest.fixture(scope="module")
def c5(gpu_available):
    from psi4_amd.molecule import water_cluster
    from psi4_amd.scf import RHF
    mol = water_cluster((4, 4, 3), seed=0)
    announce("(H2O)_48 plain FP64 RHF (C5)")
    solver = RHF(mol)
    res = solver.compute_energy()
    print(f"  E {res.energy:.12f} in {res.iterations} iterations, {res.timings}", flush=True)
    assert res.J.shape == (1152, 1152) and solver.nocc == 240
    return mol, solver, res


def test_c5_plain_energy_vs_oracle_contractions(c5):
    from oracle.adjoint import energy
    from psi4_amd import ints
    mol, solver, res = c5
    jk = solver.jk()
    naux = jk.info()["naux"]
    assert naux == 5568
    C = np.ascontiguousarray(res.C[:, :solver.nocc])
    announce(f"(H2O)_48 oracle energy over Q-slices of {SLICE}")
    e2 = 0.0
    for q0 in range(0, naux, SLICE):
        B = jk.get_B_dense(q0, min(naux, q0 + SLICE))
        e2 += energy(B, C)
        del B
    D = C @ C.T
    H = ints.kinetic(solver.basis) + ints.potential(solver.basis, mol)
    E_oracle = mol.nuclear_repulsion_energy() + 2.0 * float(np.sum(D * H)) + e2
    print(f"  E device {res.energy:.12f}  oracle contractions {E_oracle:.12f}  "
          f"dE {abs(res.energy - E_oracle):.2e}", flush=True)
    assert abs(res.energy - E_oracle) < 1e-8


def test_c5_incremental_mixed_vs_plain(c5):
    from psi4_amd.scf import RHF
    from helpers import per_build_checker
    mol, solver, res = c5
    announce("(H2O)_48 incremental BF16x3 RHF with FP64 finishing iterations")
    jk = solver.jk()
    mixed = RHF(mol, jk=jk, mixed_precision="bf16x3")
    worst = []
    mixed.jk_callback = per_build_checker(jk, worst)
    rm = mixed.compute_energy()
    kinds = [h["build"] for h in rm.history]
    print(f"  E mixed {rm.energy:.12f} plain {res.energy:.12f} dE {abs(rm.energy - res.energy):.2e}; "
          f"{kinds.count('bf16x3')} mixed builds, worst |dE_build| {max(worst):.2e}", flush=True)
    assert "bf16x3" in kinds and kinds[-1] == "fp64"
    assert abs(rm.energy - res.energy) < 1e-8
    assert max(worst) <= 1e-6


def test_c5_adjoint_identities(c5):
    mol, solver, res = c5
    jk = solver.jk()
    naux = jk.info()["naux"]
    C = np.ascontiguousarray(res.C[:, :solver.nocc])
    announce("(H2O)_48 device adjoint: <C, dE/dC> = 4E, sum over Q-slices of <B, dE/dB> = 2E")
    E, dC, _ = jk.energy_adjoint(C, False)
    ec = float(np.vdot(C, dC))
    eb = 0.0
    try:
        for q0 in range(0, naux, SLICE):
            q1 = min(naux, q0 + SLICE)
            jk.set_aux_range(q0, q1)
            Es, _, G = jk.energy_adjoint(C, True)
            assert Es == E                                   # E always covers every Q
            eb += float(np.vdot(jk.get_B_dense(q0, q1), G))
            del G
    finally:
        jk.set_aux_range(0, naux)
    print(f"  E {E:.12f}  <C,dC>/4 {ec / 4:.12f}  <B,G>/2 {eb / 2:.12f}", flush=True)
    assert abs(ec - 4 * E) < 1e-12 * abs(E)
    assert abs(eb - 2 * E) < 1e-12 * abs(E)
