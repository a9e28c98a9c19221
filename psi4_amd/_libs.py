"""Locating and loading the in-tree native libraries.

The libraries are built in-tree by ``psi4_amd.build`` (``__graft_entry__.build()``)
into ``psi4_amd/lib``.  Loading fails loudly when a library is missing: there is
no fallback implementation of any entry point.
"""
from __future__ import annotations

import ctypes
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
INTS_LIB = os.path.join(LIB_DIR, "libp4a_ints.so")
DFJK_LIB = os.path.join(LIB_DIR, "libpsi4amd_dfjk.so")

_cache = {}


def _load(path):
    if path not in _cache:
        if not os.path.exists(path):
            raise ImportError(f"native library {path} is not built; run __graft_entry__.build()")
        _cache[path] = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    return _cache[path]


def load_ints_lib():
    lib = _load(INTS_LIB)
    if not getattr(lib, "_p4a_sigs", False):
        lib.p4a_ints_last_error.restype = ctypes.c_char_p
        lib._p4a_sigs = True
    return lib


def load_dfjk_lib():
    return _load(DFJK_LIB)
