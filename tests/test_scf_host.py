"""Host logic of the SCF drivers (CPU): the orthogonalizer of psi4's S_ORTHOGONALIZATION AUTO
(libmints/orthog.cc:58-77, 176-235, 330-360, S_TOLERANCE 1e-7)."""
import numpy as np
import torch

import psi4_amd.molecule as M
from psi4_amd import ints
from psi4_amd.basisset import BasisSet
from psi4_amd.scf import orthogonalizer


def _overlap(mol, basis="cc-pvdz"):
    return torch.from_numpy(ints.overlap(BasisSet.build(mol, basis)))


def test_symmetric_branch_is_s_inverse_sqrt():
    """Well-conditioned S: X = S^-1/2 (symmetric, square), X^T S X = 1."""
    mol = M.Molecule.from_string("units bohr\nO 0 0 0\nH 0 1.43 1.1\nH 0 -1.43 1.1")
    S = _overlap(mol)
    X = orthogonalizer(S)
    n = S.shape[0]
    assert X.shape == (n, n)
    assert torch.allclose(X, X.T, atol=1e-13)
    assert torch.allclose(X.T @ S @ X, torch.eye(n, dtype=S.dtype), atol=1e-12)
    # S^-1/2 is the unique symmetric positive-definite inverse square root: X S X = 1 and X > 0
    assert float(torch.linalg.eigvalsh(X)[0]) > 0.0


def test_canonical_branch_drops_near_dependencies():
    """Two water molecules 1e-4 bohr apart make S nearly singular: the canonical branch keeps
    the eigen-directions of the normalised S above 1e-7, nmo < nbf, and X^T S X = 1 on them;
    the dropped directions are the near-null space of S."""
    base = np.array([[0.0, 0.0, 0.0], [0.0, 1.43, 1.1], [0.0, -1.43, 1.1]])
    xyz = np.vstack([base, base + 1e-4])
    mol = M.Molecule(["O", "H", "H", "O", "H", "H"], xyz)
    S = _overlap(mol)
    n = S.shape[0]
    d = 1.0 / torch.sqrt(torch.diagonal(S))
    s = torch.linalg.eigvalsh(S * d[:, None] * d[None, :])
    assert float(s[0]) < 1e-7                       # the test needs the near-singular case
    X = orthogonalizer(S)
    nmo = X.shape[1]
    assert nmo == int((s > 1e-7).sum()) and nmo < n
    assert torch.allclose(X.T @ S @ X, torch.eye(nmo, dtype=S.dtype), atol=1e-9)
    # everything S-orthogonal to span(X) has tiny norm in the S metric
    P = X @ X.T @ S                                 # S-orthogonal projector onto span(X)
    R = torch.eye(n, dtype=S.dtype) - P
    w = torch.linalg.eigvalsh(R.T @ S @ R)
    assert float(w[-1]) < 1e-6


def test_rank_deficient_overlap_raises():
    mol = M.Molecule.from_string("units bohr\nO 0 0 0\nH 0 1.43 1.1\nH 0 -1.43 1.1")
    S = _overlap(mol)
    Sd = torch.block_diag(S, S)
    Sd[: S.shape[0], S.shape[0]:] = S
    Sd[S.shape[0]:, : S.shape[0]] = S               # exactly duplicated functions
    try:
        orthogonalizer(Sd)
    except RuntimeError as e:
        assert "overcomplete" in str(e)
    else:
        raise AssertionError("a singular overlap must raise")
