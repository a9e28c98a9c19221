"""Adjoint of E = tr(D(2J - K)) on the GPU (dfjk_energy_adjoint[_device]) and its
torch.autograd.Function, against the oracle and central differences.

Workload: a seeded (H2O)_16 cluster, cc-pVDZ/cc-pVDZ-jkfit (nbf 384, nocc 80),
at the converged FP64 RHF orbitals.  The oracle (oracle/adjoint.py, pinned in
tests/test_oracle_pins.py) runs on the device's own fitted tensor B, so the
comparison isolates the adjoint kernels from the metric factorisation.

Tolerances (FP64): E to 1e-12 relative; dE/dC and dE/dB to 1e-12 of their
largest element; central differences (step 1e-5; E is exactly quadratic in B
and quartic in C) to 2e-6 of max(1, |derivative|).
"""
import numpy as np
import pytest

from conftest import announce
from helpers import maxabs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def h2o16(gpu_available):
    from psi4_amd.molecule import water_cluster
    from psi4_amd.scf import RHF
    mol = water_cluster((4, 2, 2), seed=1)
    announce("(H2O)_16 FP64 RHF on the GPU JK (orbitals for the adjoint)")
    solver = RHF(mol)
    res = solver.compute_energy()
    jk = solver.jk()
    naux = jk.info()["naux"]
    B = jk.get_B_dense(0, naux)
    C = np.ascontiguousarray(res.C[:, :solver.nocc])
    return jk, B, C, solver


def _rel(a, b):
    return maxabs(a, b) / max(1.0, float(np.abs(b).max()))


def test_adjoint_vs_oracle_h2o16(h2o16):
    from oracle.adjoint import energy_adjoint
    jk, B, C, _ = h2o16
    naux = B.shape[0]
    assert C.shape == (384, 80)
    announce("adjoint host entry, Q slice [40, 296) vs oracle")
    jk.set_aux_range(40, 296)
    try:
        E, dC, G = jk.energy_adjoint(C, True)
    finally:
        jk.set_aux_range(0, naux)
    Eo, dCo, Go = energy_adjoint(B, C, 40, 296)
    print(f"  E {E:.12f} oracle {Eo:.12f}  dE/dC rel {_rel(dC, dCo):.2e}  Gamma rel {_rel(G, Go):.2e}", flush=True)
    assert abs(E - Eo) < 1e-12 * abs(Eo)
    assert _rel(dC, dCo) < 1e-12
    assert G.shape == (256, 384, 384)
    assert _rel(G, Go) < 1e-12


def test_adjoint_homogeneity_full_range(h2o16):
    """E is homogeneous of degree 2 in B and 4 in C: <B, dE/dB> = 2E, <C, dE/dC> = 4E."""
    jk, B, C, _ = h2o16
    announce("adjoint host entry, all Q: Euler identities")
    E, dC, G = jk.energy_adjoint(C, True)
    assert G.shape == B.shape
    eb = float(np.vdot(B, G))
    ec = float(np.vdot(C, dC))
    print(f"  E {E:.12f}  <B,G>/2 {eb / 2:.12f}  <C,dC>/4 {ec / 4:.12f}", flush=True)
    assert abs(eb - 2 * E) < 1e-11 * abs(E)
    assert abs(ec - 4 * E) < 1e-11 * abs(E)
    assert np.abs(G - G.transpose(0, 2, 1)).max() < 1e-12 * np.abs(G).max()


def test_autograd_vs_central_differences_h2o16(h2o16):
    import torch
    from oracle.dfhelper import MemDFJKOracle
    from psi4_amd.autograd import scf_energy_2e
    jk, B, C, solver = h2o16
    dev = torch.device("cuda:0")
    announce("autograd: E, dE/dC, dE/dB through torch.autograd.Function")
    Ct = torch.from_numpy(C).to(dev).requires_grad_(True)
    Bt = torch.from_numpy(B).to(dev).requires_grad_(True)
    E = scf_energy_2e(jk, Ct, Bt)
    E.backward()
    gC = Ct.grad.cpu().numpy()
    gB = Bt.grad

    # 4(2J - K)C from the oracle's MemDFJK contractions on the same B
    _, J, K = MemDFJKOracle(solver.basis, solver.aux, B=B).compute([C])
    ref = 4.0 * (2.0 * J[0] - K[0]) @ C
    print(f"  E {float(E.detach()):.12f}  dE/dC vs 4(2J-K)C rel {_rel(gC, ref):.2e}", flush=True)
    assert _rel(gC, ref) < 1e-12

    rng = np.random.default_rng(20260116)
    h = 1e-5
    worst = 0.0
    announce("central differences on 32 entries of C")
    Cw = Ct.detach().clone()
    for _ in range(32):
        m, i = int(rng.integers(C.shape[0])), int(rng.integers(C.shape[1]))
        Cw[m, i] += h
        ep = float(scf_energy_2e(jk, Cw))
        Cw[m, i] -= 2 * h
        em = float(scf_energy_2e(jk, Cw))
        Cw[m, i] += h
        fd = (ep - em) / (2 * h)
        err = abs(fd - gC[m, i]) / max(1.0, abs(gC[m, i]))
        worst = max(worst, err)
        assert err < 2e-6, (m, i, fd, gC[m, i])
    print(f"  worst C central-difference error {worst:.2e}", flush=True)

    announce("central differences on 32 stored entries of B (symmetric pairs)")
    Bw = Bt.detach().clone()
    Cd = Ct.detach()
    worst = 0.0
    n = C.shape[0]
    done = 0
    while done < 32:
        q, m, k = int(rng.integers(B.shape[0])), int(rng.integers(n)), int(rng.integers(n))
        if B[q, m, k] == 0.0:
            continue  # screened pair: not stored
        pair = [(m, k)] if m == k else [(m, k), (k, m)]
        for a, b in pair:
            Bw[q, a, b] += h
        ep = float(scf_energy_2e(jk, Cd, Bw))
        for a, b in pair:
            Bw[q, a, b] -= 2 * h
        em = float(scf_energy_2e(jk, Cd, Bw))
        for a, b in pair:
            Bw[q, a, b] += h
        fd = (ep - em) / (2 * h)
        an = float(sum(gB[q, a, b] for a, b in pair))
        err = abs(fd - an) / max(1.0, abs(an))
        worst = max(worst, err)
        assert err < 2e-6, (q, m, k, fd, an)
        done += 1
    print(f"  worst B central-difference error {worst:.2e}", flush=True)


def test_autograd_leaves_the_jk_unchanged(h2o16):
    """scf_energy_2e with a caller's B (forward and backward) does not touch the JK object:
    a compute() after it equals one before it, bit for bit, and its aux range is kept."""
    import torch
    from psi4_amd.autograd import scf_energy_2e
    jk, B, C, _ = h2o16
    dev = torch.device("cuda:0")
    announce("autograd side effects: compute() before and after scf_energy_2e(jk, C, B')")
    naux = B.shape[0]
    jk.C_clear()
    jk.C_left_add(C)
    jk.compute()
    J0, K0 = jk.J()[0].copy(), jk.K()[0].copy()
    jk.set_aux_range(10, naux - 10)
    Bp = torch.from_numpy(1.01 * B).to(dev).requires_grad_(True)
    Ct = torch.from_numpy(C).to(dev).requires_grad_(True)
    E = scf_energy_2e(jk, Ct, Bp)
    E.backward()
    assert jk.info()["aux_range"] == (10, naux - 10)
    jk.set_aux_range(0, naux)
    jk.compute()
    assert np.array_equal(jk.J()[0], J0) and np.array_equal(jk.K()[0], K0)
    assert np.array_equal(jk.get_B_dense(0, 4), B[:4])


@pytest.mark.parametrize("nocc", [72, 330])
def test_adjoint_tensor_blocks_and_panels(gpu_available, nocc):
    """Synthetic tensor, Q processed in several blocks (small memory budget), one and two
    column panels of the half-transform, against the oracle."""
    from oracle.adjoint import energy_adjoint
    from psi4_amd.core import JK
    nbf, naux = 200, 150
    announce(f"adjoint on a synthetic tensor nbf {nbf} naux {naux} nocc {nocc}, blocked")
    jk = JK.build_tensor(nbf, naux)
    jk.initialize()
    jk.fill_synthetic(77)
    est = jk.device_footprint()
    jk.set_memory(int(est + 16 * (256 * 384 + nocc * nocc + nbf * nocc + nbf * nbf)))
    B = jk.get_B_dense(0, naux)
    C = np.random.default_rng(nocc).standard_normal((nbf, nocc)) / np.sqrt(nbf)
    E, dC, G = jk.energy_adjoint(C, True)
    Eo, dCo, Go = energy_adjoint(B, C)
    print(f"  E {E:.12f} oracle {Eo:.12f}  dE/dC rel {_rel(dC, dCo):.2e}  Gamma rel {_rel(G, Go):.2e}", flush=True)
    assert abs(E - Eo) < 1e-12 * abs(Eo)
    assert _rel(dC, dCo) < 1e-12
    assert _rel(G, Go) < 1e-12
