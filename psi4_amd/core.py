"""psi4-style entry points over the native JK object.

``JK.build`` restates psi4's ``core.JK.build`` (psi4/driver/p4util/python_helpers.py:442-509)
plus the option mapping of ``_set_dfjk_options`` (libfock/jk.cc:58-68) and the
SCREENING checks of ``JK::build_JK`` (jk.cc:84-109) for the in-core DF
algorithm.  The returned object is ``psi4_amd._core.JK`` (pybind11 over the C
ABI in include/psi4amd_dfjk.h); its method names follow export_fock.cc:49-97.
"""
from __future__ import annotations

from ._libs import load_dfjk_lib, load_ints_lib

load_ints_lib()
load_dfjk_lib()
from . import _core  # noqa: E402  (the native module; importing it fails loudly if not built)

PREC_FP64 = _core.PREC_FP64
PREC_FP32 = _core.PREC_FP32
PREC_BF16X3 = _core.PREC_BF16X3
FIT_CHOLESKY = _core.FIT_CHOLESKY
FIT_EIGEN = _core.FIT_EIGEN

_SUPPORTED = ("DF", "MEM_DF")


def _jk_build(orbital_basis, aux=None, jk_type=None, do_wK=None, memory=None, *, screening="SCHWARZ",
              ints_tolerance=None, df_fitting_condition=1e-10, df_basis_scf=None):
    """Construct an uninitialized in-core DF JK object (psi4 ``core.JK.build``).

    Parameters follow psi4; option values that psi4 reads from its global
    options (SCREENING, INTS_TOLERANCE, DF_FITTING_CONDITION, DF_BASIS_SCF)
    are keyword arguments here.
    """
    from .basisset import BasisSet
    jk_type = (jk_type or "MEM_DF").upper()
    if jk_type not in _SUPPORTED:
        raise RuntimeError(f"JK type {jk_type} is not provided by this engine (only MEM_DF)")
    screening = screening.upper()
    if screening == "DENSITY":                       # jk.cc:84-109
        raise RuntimeError("SCREENING=DENSITY has not been implemented for MEM_DF.")
    if aux is None:                                  # python_helpers.py:495-498
        name = df_basis_scf or (orbital_basis.name + "-jkfit")
        aux = BasisSet.build(orbital_basis.molecule, name, puream=orbital_basis.puream)
    if do_wK is None or memory is None:
        jk = _core.JK.build_JK(orbital_basis, aux)
    else:
        jk = _core.JK.build_JK(orbital_basis, aux, bool(do_wK), int(memory))
    # _set_dfjk_options (jk.cc:58-68): the cutoff is set only if INTS_TOLERANCE was
    # given or SCREENING is NONE; otherwise the default 1e-12 stays.
    if screening == "NONE":
        jk.set_cutoff(0.0)
    elif ints_tolerance is not None:
        jk.set_cutoff(float(ints_tolerance))
    jk.set_condition(float(df_fitting_condition))
    return jk


class JK:
    """Namespace mirroring ``psi4.core.JK``: ``JK.build(...)`` and ``JK.build_JK(...)``."""
    build = staticmethod(_jk_build)
    build_JK = staticmethod(_core.JK.build_JK)
    build_tensor = staticmethod(_core.JK.build_tensor)
