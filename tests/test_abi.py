"""C ABI and host-side logic (CPU, no GPU): the libraries load, every entry
point declared in include/*.h is exported, and the JK object's host-side
checks raise like psi4's."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", text, flags=re.M)))


@pytest.mark.parametrize("header,lib", [("psi4amd_ints.h", "libp4a_ints.so"),
                                        ("psi4amd_dfjk.h", "libpsi4amd_dfjk.so")])
def test_every_declared_symbol_is_exported(header, lib):
    names = _declared(header)
    assert len(names) > 5
    so = ctypes.CDLL(os.path.join(ROOT, "psi4_amd", "lib", lib))
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def _dfjk():
    from psi4_amd._libs import load_dfjk_lib
    lib = load_dfjk_lib()
    lib.dfjk_last_error.restype = ctypes.c_char_p
    return lib


def test_c_abi_status_codes_without_device():
    lib = _dfjk()
    h = ctypes.c_void_p()
    assert lib.dfjk_create_tensor(ctypes.byref(h), 0, 10) == 1          # DFJK_ERR_ARG
    assert b"bad sizes" in lib.dfjk_last_error()
    assert lib.dfjk_create_tensor(ctypes.byref(h), 40, 10) == 0
    assert lib.dfjk_set_precision(h, 7) == 1
    assert lib.dfjk_set_cutoff(h, ctypes.c_double(-1.0)) == 1
    one = (ctypes.c_int * 1)(4)
    assert lib.dfjk_compute(h, 1, None, None, one, None, None, None, 0) == 2   # DFJK_ERR_STATE
    assert b"initialize" in lib.dfjk_last_error()
    est = ctypes.c_size_t(0)
    assert lib.dfjk_memory_estimate(h, ctypes.byref(est)) == 0
    # psi4's get_core_size with every pair significant, one thread, Qshell_max 1:
    # naux*nbf^2 + naux^2 + nbf^2 + 3*nbf^2
    assert est.value == 10 * 1600 + 100 + 1600 + 3 * 1600
    assert lib.dfjk_device_footprint(h, ctypes.byref(est)) == 0
    # 40 -> 2 tiles of 32 -> 3 stored tile pairs of 32x32 per Q
    assert est.value >= 10 * 3 * 1024
    assert lib.dfjk_set_pair_form(h, 5) == 1
    assert lib.dfjk_set_pair_form(h, 2) == 0
    q0, q1 = ctypes.c_int(-1), ctypes.c_int(-1)
    assert lib.dfjk_get_aux_range(h, ctypes.byref(q0), ctypes.byref(q1)) == 0
    assert (q0.value, q1.value) == (0, 10)
    fit, kept, cond = ctypes.c_int(-1), ctypes.c_int(-1), ctypes.c_double(0)
    assert lib.dfjk_get_fit_info(h, ctypes.byref(fit), ctypes.byref(cond), ctypes.byref(kept)) == 0
    assert (fit.value, cond.value, kept.value) == (0, 1e-10, 0)
    assert lib.dfjk_destroy(h) == 0


def test_jk_build_host_logic():
    from psi4_amd.basisset import BasisSet
    from psi4_amd.core import JK
    from psi4_amd.molecule import water_cluster
    mol = water_cluster((2, 2, 1), seed=0)
    bs = BasisSet.build(mol, "cc-pvdz")
    jk = JK.build(bs)
    assert jk.name() == "MemDFJK"
    info = jk.info()
    assert info["nbf"] == 96 and info["naux"] == 464
    assert jk.memory_estimate() > 464 * 96 * 96 // 2
    with pytest.raises(RuntimeError, match="DENSITY"):
        JK.build(bs, screening="DENSITY")
    with pytest.raises(RuntimeError, match="only MEM_DF"):
        JK.build(bs, jk_type="DIRECT")
    est = jk.memory_estimate()
    foot = jk.device_footprint()
    jk.set_do_wK(True)                  # the attenuated tensor doubles the resident device part
    assert jk.device_footprint() - foot == 464 * jk.info()["ntile_pairs"] * 1024
    assert jk.memory_estimate() == est  # psi4's MemDFJK estimate does not count wK (MemDFJK.cc:65-69)
    from psi4_amd.core import PREC_BF16X3, PREC_FP32, PREC_FP64
    bq = 464 * jk.info()["ntile_pairs"] * 1024   # B's doubles
    wk = jk.device_footprint()
    jk.set_precision(PREC_FP32)         # the FP32 copy: half of B's FP64 bytes
    assert jk.device_footprint() - wk == bq // 2
    jk.set_precision(PREC_BF16X3)       # and its tile-transposed twin
    assert jk.device_footprint() - wk == bq
    jk.set_precision(PREC_FP64)
    assert jk.device_footprint() == wk
    with pytest.raises(RuntimeError, match="omega"):
        jk.set_omega(-1.0)
    jk.set_wcombine(True)               # combined exchange: argument checks of the weights
    assert jk.get_wcombine()
    with pytest.raises(RuntimeError, match="omega_alpha"):
        jk.set_omega_alpha(float("nan"))
    with pytest.raises(RuntimeError, match="omega_beta"):
        jk.set_omega_beta(float("inf"))
    jk.set_omega_alpha(0.25)
    jk.set_omega_beta(0.75)
    assert (jk.get_omega_alpha(), jk.get_omega_beta()) == (0.25, 0.75)
    jk.set_wcombine(False)
    jk.set_do_wK(False)
    with pytest.raises(RuntimeError, match="initialize"):
        jk.C_left_add(np.zeros((96, 5)))
        jk.compute()
    s = jk.print_header()
    assert "MemDFJK" in s and "Schwarz Cutoff" in s


def test_gbs_parser_and_cluster_generator():
    from psi4_amd.basisset import parse_gbs
    from psi4_amd.molecule import water_cluster
    pure, atoms = parse_gbs(os.path.join(ROOT, "psi4_amd", "basis", "cc-pvdz.gbs"))
    assert pure and [s[0] for s in atoms["O"]] == [0, 0, 0, 1, 1, 2]
    m1 = water_cluster((3, 4, 2), seed=0)
    m2 = water_cluster((3, 4, 2), seed=0)
    assert m1.natom == 72 and np.array_equal(m1.xyz, m2.xyz)
    # rigid waters: O-H 0.9572 A, HOH 104.52 deg
    from psi4_amd.molecule import BOHR2ANGSTROM
    x = m1.xyz * BOHR2ANGSTROM
    for w in range(24):
        o, h1, h2 = x[3 * w], x[3 * w + 1], x[3 * w + 2]
        assert abs(np.linalg.norm(h1 - o) - 0.9572) < 1e-12
        c = np.dot(h1 - o, h2 - o) / (np.linalg.norm(h1 - o) * np.linalg.norm(h2 - o))
        assert abs(np.degrees(np.arccos(c)) - 104.52) < 1e-9


def test_product_schwarz_count_and_memory_estimate_ar5():
    """The product's own screening and memory estimate through the ABI, pinned by psi4's
    tests/pytests/test_jkmemory.py:37-45: Ar5 cc-pVDZ/cc-pVDZ-jkfit with one thread gives
    1962 significant function pairs of 8100 and a MEM_DF estimate of 1,590,520 doubles."""
    import psi4_amd.molecule as M
    from psi4_amd.basisset import BasisSet
    from psi4_amd.core import JK
    mol = M.Molecule.from_string("Ar 0 0 0\nAr 0 0 5\nAr 0 0 15\nAr 0 0 25\nAr 0 0 35")
    prim, aux = BasisSet.build(mol, "cc-pvdz"), BasisSet.build(mol, "cc-pvdz-jkfit")
    jk = JK.build(prim, aux)
    jk.set_omp_nthread(1)                      # psi4 -n1
    assert jk.memory_estimate() == 1590520
    info = jk.info()
    assert (info["nbf"], info["nsig_pairs"]) == (90, 1962)
    mask = jk.pair_mask()
    assert mask.shape == (90, 90) and int(mask.sum()) == 1962
    assert np.array_equal(mask, mask.T)


@pytest.mark.parametrize("grid,seed,cutoff", [((2, 2, 1), 0, 1e-12), ((3, 4, 2), 0, 1e-12), ((2, 2, 1), 0, 1e-6)])
def test_product_schwarz_mask_equals_oracle(grid, seed, cutoff):
    """The product's Schwarz function-pair mask is the oracle's prepare_sparsity mask
    (dfhelper.cc:299-420) bit for bit, on (H2O)_4 and (H2O)_24 and with a coarse cutoff."""
    from oracle.dfhelper import prepare_sparsity
    from psi4_amd import ints
    from psi4_amd.basisset import BasisSet
    from psi4_amd.core import JK
    from psi4_amd.molecule import water_cluster
    mol = water_cluster(grid, seed=seed)
    prim, aux = BasisSet.build(mol, "cc-pvdz"), BasisSet.build(mol, "cc-pvdz-jkfit")
    jk = JK.build(prim, aux, ints_tolerance=cutoff)
    mask = jk.pair_mask()
    fm, sm, mx = ints.schwarz(prim)
    sp = prepare_sparsity(fm, sm, mx, cutoff, aux.nbf())
    ref = (np.asarray(sp.fun_mask) != 0).astype(np.uint8)
    assert mask.shape == ref.shape
    assert np.array_equal(mask, ref)
    assert jk.info()["nsig_pairs"] == sp.nsig_pairs()
    assert jk.info()["cutoff"] == cutoff


def test_schwarz_values_through_the_abi_without_device():
    """dfjk_get_schwarz on a machine without a GPU returns the host engine's screening integrals
    (psi4's TwoBodyAOInt function_pair_values_ / shell_pair_value / max_integral,
    libmints/twobody.h:232-238): |(mn|mn)| symmetric, shell maxima the block maxima, and the
    same values as the engine's own schwarz call."""
    from psi4_amd import ints
    from psi4_amd.basisset import BasisSet
    from psi4_amd.core import JK
    from psi4_amd.molecule import water_cluster
    mol = water_cluster((2, 1, 1), seed=0)
    prim, aux = BasisSet.build(mol, "cc-pvdz"), BasisSet.build(mol, "cc-pvdz-jkfit")
    jk = JK.build(prim, aux)
    fun, shell, mx, on_dev = jk.schwarz_values()
    assert not on_dev
    fm, sm, mxr = ints.schwarz(prim)
    assert np.array_equal(fun, fm) and np.array_equal(shell, sm) and mx == mxr
    assert np.array_equal(fun, fun.T) and mx == fun.max()
    off = np.concatenate([prim.shell_offsets(), [prim.nbf()]])
    for M in range(prim.nshell):
        for N in range(prim.nshell):
            assert shell[M, N] == fun[off[M]:off[M + 1], off[N]:off[N + 1]].max()
