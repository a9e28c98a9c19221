"""ORACLE (test infrastructure): the seeded synthetic tensor-level workload.

B_Qmn = s_Q * exp(-|m-n|/64) * (G_Qmn + G_Qnm)/2, s_Q = (Q+1)^-1/2, G ~ N(0,1)
from Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) keyed by the 64-bit
seed with counter (m, n, Q, 0); the two 53-bit uniforms of each draw feed a
Box-Muller transform.  The HIP generator in psi4_amd implements the same
definition; this restatement checks it (Philox known-answer vectors from the
Random123 distribution pin the counter-based generator).
"""
from __future__ import annotations

import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10 on uint32 arrays; returns four uint32 arrays."""
    c0, c1, c2, c3 = (np.asarray(x, dtype=np.uint32) for x in (c0, c1, c2, c3))
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    with np.errstate(over="ignore"):
        for _ in range(10):
            p0 = M0 * c0.astype(np.uint64)
            p1 = M1 * c2.astype(np.uint64)
            hi0, lo0 = (p0 >> np.uint64(32)).astype(np.uint32), (p0 & MASK32).astype(np.uint32)
            hi1, lo1 = (p1 >> np.uint64(32)).astype(np.uint32), (p1 & MASK32).astype(np.uint32)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
            k0 = np.uint32((int(k0) + int(W0)) & 0xFFFFFFFF)
            k1 = np.uint32((int(k1) + int(W1)) & 0xFFFFFFFF)
    return c0, c1, c2, c3


def normals(m, n, Q, seed):
    x0, x1, x2, x3 = philox4x32_10(m, n, Q, np.zeros_like(np.asarray(m, dtype=np.uint32)),
                                   seed & 0xFFFFFFFF, seed >> 32)
    i1 = (x0.astype(np.uint64) << np.uint64(21)) | (x1.astype(np.uint64) >> np.uint64(11))
    i2 = (x2.astype(np.uint64) << np.uint64(21)) | (x3.astype(np.uint64) >> np.uint64(11))
    u1 = (i1 + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = i2.astype(np.float64) * 2.0 ** -53
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


def synthetic_B(nbf, q0, q1, seed):
    """Dense B_Q for Q in [q0, q1): array [q1-q0, nbf, nbf]."""
    m, n = np.meshgrid(np.arange(nbf, dtype=np.uint32), np.arange(nbf, dtype=np.uint32), indexing="ij")
    decay = np.exp(-np.arange(nbf, dtype=np.float64) / 64.0)
    D = decay[np.abs(m.astype(np.int64) - n.astype(np.int64))]
    out = np.empty((q1 - q0, nbf, nbf))
    for Q in range(q0, q1):
        Qa = np.full_like(m, Q)
        g = 0.5 * (normals(m, n, Qa, seed) + normals(n, m, Qa, seed))
        out[Q - q0] = (1.0 / np.sqrt(Q + 1.0)) * D * g
    return out

