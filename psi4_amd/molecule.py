"""Molecules and the seeded water-cluster generator.

Coordinates are held in bohr.  The Angstrom->bohr factor is psi4's
``pc_bohr2angstroms = 0.52917721067`` (psi4/include/psi4/physconst.h:402).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

BOHR2ANGSTROM = 0.52917721067
SYMBOL_Z = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10, "Ar": 18}


@dataclass
class Molecule:
    symbols: list
    xyz: np.ndarray  # (natom, 3) bohr
    charge: int = 0
    multiplicity: int = 1
    name: str = "mol"
    extra: dict = field(default_factory=dict)

    @classmethod
    def from_angstrom(cls, symbols, xyz_ang, **kw):
        return cls(list(symbols), np.asarray(xyz_ang, dtype=float) / BOHR2ANGSTROM, **kw)

    @classmethod
    def from_string(cls, text: str, units: str = "angstrom", **kw):
        """Plain 'Sym x y z' lines (the Cartesian subset of psi4's geometry syntax)."""
        syms, xyz = [], []
        for line in text.strip().splitlines():
            tok = line.split()
            if not tok:
                continue
            if tok[0].lower() == "units":
                units = tok[1].lower()
                continue
            if tok[0].lower() in ("no_com", "no_reorient", "symmetry"):
                continue
            syms.append(tok[0].capitalize())
            xyz.append([float(t) for t in tok[1:4]])
        xyz = np.array(xyz, dtype=float)
        if units.startswith("ang"):
            xyz = xyz / BOHR2ANGSTROM
        return cls(syms, xyz, **kw)

    @property
    def natom(self):
        return len(self.symbols)

    @property
    def Z(self):
        return np.array([SYMBOL_Z[s] for s in self.symbols], dtype=float)

    def nelectron(self):
        return int(self.Z.sum()) - self.charge

    def nuclear_repulsion_energy(self):
        Z, R = self.Z, self.xyz
        e = 0.0
        for i in range(self.natom):
            for j in range(i):
                e += Z[i] * Z[j] / np.linalg.norm(R[i] - R[j])
        return e

    def nuclear_repulsion_energy_deriv1(self):
        """d E_nuc / d R as [natom, 3] (libmints/molecule.cc nuclear_repulsion_energy_deriv1)."""
        Z, R = self.Z, self.xyz
        g = np.zeros((self.natom, 3))
        for i in range(self.natom):
            for j in range(i):
                d = R[i] - R[j]
                f = Z[i] * Z[j] * d / np.linalg.norm(d) ** 3
                g[i] -= f
                g[j] += f
        return g


def water_zmat(r_oh: float, theta_deg: float) -> Molecule:
    """H2O from the Z-matrix  O; H 1 r; H 1 r 2 theta  (Angstrom, degrees)."""
    t = math.radians(theta_deg)
    xyz = [[0.0, 0.0, 0.0], [0.0, 0.0, r_oh], [r_oh * math.sin(t), 0.0, r_oh * math.cos(t)]]
    return Molecule.from_angstrom(["O", "H", "H"], xyz, name="h2o")


def _random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform SO(3) rotation from a uniform unit quaternion (Shoemake 1992)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([
        math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
        math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
        math.sqrt(u1) * math.sin(2 * math.pi * u3),
        math.sqrt(u1) * math.cos(2 * math.pi * u3),
    ])
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def water_cluster(grid=(2, 2, 1), seed: int = 0, spacing: float = 2.9, jitter: float = 0.2,
                  r_oh: float = 0.9572, theta_deg: float = 104.52) -> Molecule:
    """Seeded (H2O)_n cluster: rigid waters (O-H 0.9572 A, HOH 104.52 deg) with
    uniformly random orientations, O atoms on a grid of ``spacing`` Angstrom,
    each jittered uniformly by +-``jitter`` Angstrom along every axis."""
    rng = np.random.default_rng(seed)
    half = math.radians(theta_deg) / 2
    # body frame: O at origin, H's in the xz plane symmetric about z
    body = np.array([[0.0, 0.0, 0.0],
                     [r_oh * math.sin(half), 0.0, r_oh * math.cos(half)],
                     [-r_oh * math.sin(half), 0.0, r_oh * math.cos(half)]])
    syms, xyz = [], []
    nx, ny, nz = grid
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                o = np.array([ix, iy, iz], dtype=float) * spacing + rng.uniform(-jitter, jitter, 3)
                R = _random_rotation(rng)
                for s, p in zip("OHH", body):
                    syms.append(s)
                    xyz.append(o + R @ p)
    n = nx * ny * nz
    return Molecule.from_angstrom(syms, np.array(xyz), name=f"(H2O)_{n} seed {seed}")
