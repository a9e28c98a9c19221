"""dfjk_compute_device_noalloc through the C ABI: the same results as dfjk_compute_device,
on two calls in a row (the second with fewer columns, so the shared scratch array holds the
first call's contents), FP64 to 1e-11 max-abs like every FP64 parity test; a symmetric pair in
FP32 and the cross pair form in BF16x3 to their rounding; a system beyond the scratch array is
refused before any work is queued."""
import ctypes

import numpy as np
import pytest

from conftest import announce
from helpers import gpu_jk, maxabs, random_orbitals, water_system

pytestmark = pytest.mark.gpu

TOL = 1e-11


def _lib():
    from psi4_amd._libs import load_dfjk_lib
    lib = load_dfjk_lib()
    lib.dfjk_last_error.restype = ctypes.c_char_p
    for name in ("dfjk_compute_device", "dfjk_compute_device_noalloc"):
        getattr(lib, name).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int]
    return lib


def _ptrs(ts):
    return (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])


def _call(lib, entry, handle, Cl, Cr=None):
    """One density pair through `entry`; returns D, J, K as numpy arrays."""
    import torch
    nbf, ncol = Cl.shape
    dev = torch.device("cuda:0")
    cl = torch.from_numpy(np.ascontiguousarray(Cl)).to(dev)
    cr = torch.from_numpy(np.ascontiguousarray(Cr)).to(dev) if Cr is not None else None
    out = [torch.full((nbf, nbf), 7.0, dtype=torch.float64, device=dev) for _ in range(3)]
    ncols = (ctypes.c_int * 1)(ncol)
    keep = [_ptrs([cl]), _ptrs([cr]) if cr is not None else None, _ptrs([out[0]]), _ptrs([out[1]]),
            _ptrs([out[2]])]
    torch.cuda.synchronize()
    st = getattr(lib, entry)(ctypes.c_void_p(handle), 1, keep[0], keep[1], ncols, keep[2], keep[3], keep[4], 0)
    torch.cuda.synchronize()
    assert st == 0, lib.dfjk_last_error()
    return [o.cpu().numpy() for o in out]


@pytest.fixture(scope="module")
def h2o8(gpu_available):
    from oracle.dfhelper import MemDFJKOracle
    mol, prim, aux = water_system((2, 2, 2), seed=3)
    announce("(H2O)_8 cc-pVDZ/cc-pVDZ-jkfit: oracle + GPU initialize (no-alloc entry point)")
    return prim, MemDFJKOracle(prim, aux), gpu_jk(prim, aux)


def test_noalloc_equals_compute_device_two_calls(h2o8):
    prim, ora, jk = h2o8
    lib = _lib()
    nbf = prim.nbf()
    for i, ncol in enumerate((40, 24)):
        C = random_orbitals(nbf, ncol, seed=11 + i)
        announce(f"noalloc call {i + 1}: symmetric pair, nbf {nbf}, {ncol} columns")
        Dn, Jn, Kn = _call(lib, "dfjk_compute_device_noalloc", jk.handle(), C)
        Da, Ja, Ka = _call(lib, "dfjk_compute_device", jk.handle(), C)
        assert maxabs(Dn, Da) < 1e-13
        assert maxabs(Jn, Ja) < TOL and maxabs(Kn, Ka) < TOL
        _, Jo, Ko = ora.compute([C])
        assert maxabs(Jn, Jo[0]) < TOL and maxabs(Kn, Ko[0]) < TOL


def test_noalloc_general_pair(h2o8):
    prim, ora, jk = h2o8
    lib = _lib()
    nbf = prim.nbf()
    for i, ncol in enumerate((30, 12)):
        Cl = random_orbitals(nbf, ncol, seed=21 + i)
        Cr = random_orbitals(nbf, ncol, seed=31 + i)
        announce(f"noalloc call {i + 1}: general pair, {ncol} columns")
        _, Jn, Kn = _call(lib, "dfjk_compute_device_noalloc", jk.handle(), Cl, Cr)
        _, Ja, Ka = _call(lib, "dfjk_compute_device", jk.handle(), Cl, Cr)
        assert maxabs(Jn, Ja) < TOL and maxabs(Kn, Ka) < TOL


def test_noalloc_bf16x3_cross_form(h2o8):
    """The incremental (cross) form in BF16x3: the allocating call makes the FP32 copies of B
    (plain and tile-transposed), then both entry points agree to the mixed-precision rounding
    (different Q-blocking)."""
    prim, ora, jk = h2o8
    lib = _lib()
    nbf = prim.nbf()
    jk.set_precision(2)
    try:
        for i, half in enumerate((20, 8)):
            A = random_orbitals(nbf, half, seed=41 + i)
            V = 1e-3 * random_orbitals(nbf, half, seed=51 + i)
            Cl, Cr = np.hstack([A, V]), np.hstack([V, A])
            announce(f"noalloc call {i + 1}: BF16x3 cross pair, {2 * half} columns")
            _, Ja, Ka = _call(lib, "dfjk_compute_device", jk.handle(), Cl, Cr)
            _, Jn, Kn = _call(lib, "dfjk_compute_device_noalloc", jk.handle(), Cl, Cr)
            assert maxabs(Kn, Kn.T) == 0.0
            assert maxabs(Jn, Ja) <= 2e-5 * np.abs(Ja).max()
            assert maxabs(Kn, Ka) <= 2e-4 * np.abs(Ka).max()
    finally:
        jk.set_precision(0)


def test_noalloc_fp32_symmetric(h2o8):
    """FP32 products on the FP32 copy of B: both entry points agree to the FP32 rounding of
    differently blocked sums, on two calls in a row, the second smaller."""
    prim, ora, jk = h2o8
    lib = _lib()
    nbf = prim.nbf()
    jk.set_precision(1)
    try:
        for i, ncol in enumerate((36, 10)):
            C = random_orbitals(nbf, ncol, seed=61 + i)
            announce(f"noalloc call {i + 1}: FP32 symmetric pair, {ncol} columns")
            _, Ja, Ka = _call(lib, "dfjk_compute_device", jk.handle(), C)
            _, Jn, Kn = _call(lib, "dfjk_compute_device_noalloc", jk.handle(), C)
            assert maxabs(Kn, Kn.T) == 0.0
            assert maxabs(Jn, Ja) <= 1e-6 * np.abs(Ja).max()
            assert maxabs(Kn, Ka) <= 1e-5 * np.abs(Ka).max()
    finally:
        jk.set_precision(0)


def test_noalloc_refuses_oversized_system(gpu_available):
    """nbf 2000: D, J partials and one K split exceed the scratch array -> DFJK_ERR_MEMORY, and
    the outputs are untouched (nothing was queued)."""
    import torch
    lib = _lib()
    h = ctypes.c_void_p()
    announce("noalloc refusal: tensor handle nbf 2000, naux 64")
    assert lib.dfjk_create_tensor(ctypes.byref(h), 2000, 64) == 0
    try:
        assert lib.dfjk_initialize(h) == 0, lib.dfjk_last_error()
        dev = torch.device("cuda:0")
        C = torch.zeros((2000, 16), dtype=torch.float64, device=dev)
        out = [torch.full((2000, 2000), 7.0, dtype=torch.float64, device=dev) for _ in range(3)]
        ncols = (ctypes.c_int * 1)(16)
        torch.cuda.synchronize()
        st = lib.dfjk_compute_device_noalloc(h, 1, _ptrs([C]), None, ncols, _ptrs([out[0]]), _ptrs([out[1]]),
                                             _ptrs([out[2]]), 0)
        torch.cuda.synchronize()
        assert st == 3                                      # DFJK_ERR_MEMORY
        assert b"scratch array" in lib.dfjk_last_error()
        assert all(bool((o == 7.0).all()) for o in out)
    finally:
        lib.dfjk_destroy(h)
