"""ORACLE -- test infrastructure, not part of the product.

Adjoint of the two-electron SCF energy E = tr(D(2J - K)), D = C C^T, for a
fitted tensor B (dense [naux][nbf][nbf], symmetric in the last two indices):

    J = sum_Q B_Q d_Q,   d_Q = <B_Q, D>,   K = sum_Q B_Q D B_Q
    E = 2 sum_Q d_Q^2 - sum_Q tr(D B_Q D B_Q)
    dE/dC          = 4 (2J - K) C
    dE/dB_Qmn      = 4 d_Q D_mn - 2 (D B_Q D)_mn = 4 d_Q D_mn - 2 (C (C^T B_Q C) C^T)_mn

psi4 computes the same intermediates for its DF-SCF gradient:
  c_A = (A|mn) Dt_mn                      scfgrad/jk_grad.cc:434-437
  (A|ij) = C^T (A|mn) C                   scfgrad/jk_grad.cc:440-448
  d = (A|B)^-1 c,  J^x / K^x assembly     scfgrad/jk_grad.cc:240-262, 634-721
In the fitted basis B = M (A|mn) (M = J^-1/2 or L^-1) the fitted c and (A|ij)
become d_Q and C^T B_Q C above, and (A|pq) = (A|ij) C_pi C_qi of
scfgrad/jk_grad.cc:868+ is the D B_Q D term.

Pins (tests/test_oracle_pins.py): E equals tr(D(2J - K)) from the dfhelper
restatement (itself pinned to psi4 energies); E is exactly quadratic in B, so
central differences of E reproduce Gamma to roundoff; and dE/dC against central
differences of E in C.
"""
import numpy as np


def energy(B, C):
    """E = tr(D(2J - K)) from a dense fitted tensor, D = C C^T.

    Written for any B (symmetric or not), as psi4's J_mn = sum B_Qmn d_Q and
    K_mn = sum B_Qml D_ls B_Qsn define it, so that single elements can be perturbed."""
    D = C @ C.T
    d = np.einsum("qmn,mn->q", B, D, optimize=True)
    X = np.einsum("qmn,ni->qmi", B, C, optimize=True)          # X_Q = B_Q C
    Y = np.einsum("mi,qmj->qij", C, X, optimize=True)          # C^T B_Q C; tr(D B_Q D B_Q) = tr(Y_Q Y_Q)
    return 2.0 * float(d @ d) - float(np.einsum("qij,qji->", Y, Y, optimize=True))


def energy_adjoint(B, C, q0=0, q1=None):
    """Returns (E, dE/dC, Gamma[q0:q1]) for a dense fitted tensor B and orbitals C."""
    naux = B.shape[0]
    q1 = naux if q1 is None else q1
    D = C @ C.T
    d = np.einsum("qmn,mn->q", B, D, optimize=True)
    J = np.einsum("q,qmn->mn", d, B, optimize=True)
    X = np.einsum("qmn,ni->qmi", B, C, optimize=True)
    K = np.einsum("qmi,qni->mn", X, X, optimize=True)
    G = 2.0 * J - K
    E = float(np.sum(D * G))
    dEdC = 4.0 * G @ C
    Y = np.einsum("mi,qmj->qij", C, X[q0:q1], optimize=True)   # (Q|ij)
    DBD = np.einsum("mi,qij,nj->qmn", C, Y, C, optimize=True)
    Gam = 4.0 * d[q0:q1, None, None] * D[None] - 2.0 * DBD
    return E, dEdC, Gam
