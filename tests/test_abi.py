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
    # 2 tile pairs of 32x32 per Q: 40 -> 2 tiles -> 3 stored pairs
    assert est.value >= 10 * 3 * 1024
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
    jk.set_do_wK(True)                  # the attenuated tensor doubles the resident part
    assert jk.memory_estimate() - est == 464 * jk.info()["ntile_pairs"] * 1024
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
