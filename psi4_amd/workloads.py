"""Named workloads of the DF-JK path (sizes nbf / naux / nocc).

  h2o4       RHF energy, (H2O)_4 2x2x1 grid, cc-pVDZ/cc-pVDZ-jkfit, 96/464/20
  h2o24      plain FP64 drop-in RHF, (H2O)_24 3x4x2 grid, 576/2784/120
  synthetic  tensor-level J/K, 1500/6000/300 (seeded synthetic B and C_occ)
  h2o16      adjoint of E = tr(D(2J-K)), (H2O)_16 4x2x2 grid, 384/1856/80
  h2o48      full RHF with incremental mixed-precision Fock builds, (H2O)_48 4x4x3 grid, 1152/5568/240
"""
from __future__ import annotations

import numpy as np

WATER_GRIDS = {"h2o4": (2, 2, 1), "h2o24": (3, 4, 2), "h2o16": (4, 2, 2), "h2o48": (4, 4, 3)}
SYNTHETIC = dict(nbf=1500, naux=6000, nocc=300, seed_B=20260101, seed_C=20260102)


def synthetic_C(nbf, nocc, seed):
    """First nocc columns of the QR factor of a seeded N(0,1) nbf x nbf matrix (orthonormal)."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((nbf, nbf)))
    return np.ascontiguousarray(Q[:, :nocc])


def jk_flops(nbf, naux, nocc, symmetric=True):
    """Algorithmic flops of one J+K build (dense, no screening)."""
    nsym = nbf * (nbf + 1) // 2
    j = 4 * naux * nsym
    if symmetric:
        k_ht = 2 * naux * nocc * nbf * nbf
        k_c = naux * nocc * nbf * (nbf + 1)
    else:
        k_ht = 4 * naux * nocc * nbf * nbf
        k_c = 2 * naux * nocc * nbf * nbf
    return j, k_ht, k_c
