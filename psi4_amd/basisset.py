"""Basis sets from psi4-format .gbs files, flattened for the C ABI.

Mirrors the parts of psi4's ``BasisSet.build`` (libmints/basisset.cc) that the
DF-JK path needs: per-atom shell lists from a .gbs file, shells ordered by atom
then by file order, ``spherical``/``cartesian`` from the file header.
Normalisation of contraction coefficients happens in the native engine
(libmints/gshell.cc:60-145 semantics).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_BASIS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis")
AM = {"S": 0, "P": 1, "D": 2, "F": 3, "G": 4, "H": 5, "I": 6}


def parse_gbs(path):
    """Return (puream, {symbol: [(l, [exps], [coefs]), ...]})."""
    puream = True
    atoms = {}
    with open(path) as fh:
        lines = [ln.strip() for ln in fh]
    i = 0
    while i < len(lines):
        ln = lines[i]
        if not ln or ln.startswith("!"):
            i += 1
            continue
        low = ln.lower()
        if low in ("spherical", "cartesian"):
            puream = low == "spherical"
            i += 1
            continue
        tok = ln.split()
        if len(tok) == 2 and tok[1] == "0" and tok[0][0].isalpha():
            sym = tok[0].capitalize()
            shells = []
            i += 1
            while not lines[i].startswith("****"):
                t = lines[i].split()
                lchar, nprim = t[0].upper(), int(t[1])
                exps, coefs = [], []
                for k in range(nprim):
                    e, c = lines[i + 1 + k].replace("D", "E").replace("d", "e").split()[:2]
                    exps.append(float(e))
                    coefs.append(float(c))
                if lchar == "SP":
                    raise ValueError("SP shells are not supported")
                shells.append((AM[lchar], exps, coefs))
                i += 1 + nprim
            atoms[sym] = shells
        i += 1
    return puream, atoms


class _CBasis(ctypes.Structure):
    _fields_ = [
        ("nshell", ctypes.c_int),
        ("l", ctypes.POINTER(ctypes.c_int)),
        ("nprim", ctypes.POINTER(ctypes.c_int)),
        ("center", ctypes.POINTER(ctypes.c_double)),
        ("exps", ctypes.POINTER(ctypes.c_double)),
        ("coefs", ctypes.POINTER(ctypes.c_double)),
        ("puream", ctypes.c_int),
    ]


@dataclass
class BasisSet:
    name: str
    l: np.ndarray        # int32 [nshell]
    nprim: np.ndarray    # int32 [nshell]
    center: np.ndarray   # float64 [nshell, 3] bohr
    exps: np.ndarray     # float64 [sum nprim]
    coefs: np.ndarray    # float64 [sum nprim] (original, unnormalised)
    puream: bool
    shell_atom: np.ndarray  # int32 [nshell]
    molecule: object = None

    @classmethod
    def build(cls, mol, name: str, puream=None):
        path = os.path.join(_BASIS_DIR, name.lower() + ".gbs")
        if not os.path.exists(path):
            raise FileNotFoundError(f"basis set {name} is not bundled ({path})")
        file_pure, atoms = parse_gbs(path)
        pure = file_pure if puream is None else bool(puream)
        l, nprim, center, exps, coefs, satom = [], [], [], [], [], []
        for a, (sym, r) in enumerate(zip(mol.symbols, mol.xyz)):
            if sym not in atoms:
                raise KeyError(f"basis {name} has no entry for {sym}")
            for am, e, c in atoms[sym]:
                l.append(am)
                nprim.append(len(e))
                center.append(r)
                exps.extend(e)
                coefs.extend(c)
                satom.append(a)
        return cls(name, np.array(l, np.int32), np.array(nprim, np.int32), np.array(center, float).reshape(-1, 3),
                   np.array(exps, float), np.array(coefs, float), pure, np.array(satom, np.int32), mol)

    @property
    def nshell(self):
        return len(self.l)

    def shell_nfunc(self):
        return np.where(self.puream, 2 * self.l + 1, (self.l + 1) * (self.l + 2) // 2)

    def nbf(self):
        return int(self.shell_nfunc().sum())

    def shell_offsets(self):
        nf = self.shell_nfunc()
        return np.concatenate([[0], np.cumsum(nf)[:-1]]).astype(np.int64)

    def max_nfunc(self):
        return int(self.shell_nfunc().max())

    def as_c(self):
        """ctypes struct; keeps the arrays alive through the returned tuple."""
        arrs = (np.ascontiguousarray(self.l, np.int32), np.ascontiguousarray(self.nprim, np.int32),
                np.ascontiguousarray(self.center, np.float64), np.ascontiguousarray(self.exps, np.float64),
                np.ascontiguousarray(self.coefs, np.float64))
        s = _CBasis(self.nshell,
                    arrs[0].ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                    arrs[1].ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                    arrs[2].ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                    arrs[3].ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                    arrs[4].ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                    int(self.puream))
        return s, arrs
