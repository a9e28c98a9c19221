"""Integrals on the device against the host engine: three-index (P|mn), the metric (P|Q) and its
factor, and the Schwarz diagonal (mn|mn) of the screening.

The metric factor L^-1 (J = L L^T) from the device's blocked Cholesky and block forward
substitution gives the same fitted tensor as the host factorisation ((H2O)_16, naux 1856 in
256-row blocks with a partial last block; 1e-11 of the largest element).

The metric (P|Q) computed on the device gives the same fitted tensor as the host engine's
metric ((H2O)_16; 1e-11 of the largest element: the fit amplifies last-bit differences of
the metric by its condition number, about 1e7 here).

The fitted tensor B = L^-1 (P|mn) built from the device integrals equals the one built from
the host engine's integrals (same host metric, same device fit) to 1e-12 of the largest
element, for water and (H2O)_8 in cc-pVDZ / cc-pVDZ-jkfit (s..d | s..f) and water in
cc-pVTZ / cc-pVTZ-jkfit (s..f | s..g), and likewise the attenuated tensor of wK (omega = 0.3).
The host engine's integrals are themselves pinned by psi4's energies and gradients.

The Schwarz diagonal |(mn|mn)| from the device equals the host engine's to 1e-12 relative per
element, and the screening masks built from it are the host's bit for bit.
"""
import os
import time

import numpy as np
import pytest

import psi4_amd.molecule as M
from conftest import announce

pytestmark = pytest.mark.gpu

# host integrals imply the host metric: the three-index comparisons hold the metric on the
# host in both runs, so the two tensors differ only by their (P|mn)
HOST_METRIC = ("PSI4AMD_HOST_METRIC",)


def _tensor(mol, basis, device, omega=None, key="PSI4AMD_HOST_INTS", info=None, fixed=()):
    """B (and the attenuated tensor) with the step named by key on the device or, with
    device False, on the host; the switches in fixed are set in both runs."""
    from psi4_amd.basisset import BasisSet
    from psi4_amd.core import JK
    prim = BasisSet.build(mol, basis)
    aux = BasisSet.build(mol, basis + "-jkfit")
    keys = (key,) + tuple(fixed)
    saved = {k: os.environ.pop(k, None) for k in keys}
    if not device:
        os.environ[key] = "1"
    for k in fixed:
        os.environ[k] = "1"
    try:
        jk = JK.build(prim, aux, do_wK=bool(omega), memory=8_000_000_000)
        jk.set_print(0)
        if omega:
            jk.set_omega(omega)
        t0 = time.perf_counter()
        jk.initialize()
        dt = time.perf_counter() - t0
        naux = jk.info()["naux"]
        if info is not None:
            info.update(jk.info())
        B = jk.get_B_dense(0, naux)
        Bw = jk.get_Bw_dense(0, naux) if omega else None
    finally:
        for k, v in saved.items():
            os.environ.pop(k, None)
            if v is not None:
                os.environ[k] = v
    return B, Bw, dt


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("case", ["water-dz", "cluster8-dz", "water-tz"])
def test_device_three_index_matches_host(gpu_available, case):
    if case == "cluster8-dz":
        mol, basis = M.water_cluster((2, 2, 2), seed=9), "cc-pvdz"
    else:
        mol, basis = M.water_zmat(0.96, 104.5), ("cc-pvtz" if case == "water-tz" else "cc-pvdz")
    announce(f"device (P|mn) {case}: device integrals")
    Bd, _, td = _tensor(mol, basis, True, fixed=HOST_METRIC)
    announce(f"device (P|mn) {case}: host integrals")
    Bh, _, th = _tensor(mol, basis, False, fixed=HOST_METRIC)
    e = _rel(Bd, Bh)
    print(f"  {case}: |B_dev - B_host|/max {e:.2e}; initialize {td:.2f} s device ints, {th:.2f} s host ints", flush=True)
    assert e < 1e-12


def test_device_three_index_attenuated(gpu_available):
    mol = M.water_cluster((2, 1, 1), seed=3)
    announce("device (P|w|mn), omega 0.3, (H2O)_2 cc-pVDZ")
    Bd, Bwd, _ = _tensor(mol, "cc-pvdz", True, omega=0.3, fixed=HOST_METRIC)
    Bh, Bwh, _ = _tensor(mol, "cc-pvdz", False, omega=0.3, fixed=HOST_METRIC)
    print(f"  B {_rel(Bd, Bh):.2e}  Bw {_rel(Bwd, Bwh):.2e}", flush=True)
    assert _rel(Bd, Bh) < 1e-12 and _rel(Bwd, Bwh) < 1e-12


def test_device_metric_factor_matches_host(gpu_available):
    mol = M.water_cluster((4, 2, 2), seed=4)
    announce("metric factor (H2O)_16 cc-pVDZ-jkfit: device blocked Cholesky vs host")
    info = {}
    Bd, _, td = _tensor(mol, "cc-pvdz", True, key="PSI4AMD_HOST_FACTOR", info=info, fixed=HOST_METRIC)
    Bh, _, th = _tensor(mol, "cc-pvdz", False, key="PSI4AMD_HOST_FACTOR", fixed=HOST_METRIC)
    e = _rel(Bd, Bh)
    t = info["setup_ms"]
    print(f"  naux {info['naux']}: |B_dev - B_host|/max {e:.2e}; initialize {td:.2f} s device factor, "
          f"{th:.2f} s host factor; phases (ms) {t}", flush=True)
    assert info["naux"] == 1856
    assert e < 1e-11
    assert t["metric"] > 0 and t["metric_factor"] > 0 and t["fit"] > 0 and t["attenuated"] == 0


def test_device_metric_matches_host(gpu_available):
    """The two-centre metric (P|Q) on the device (the three-centre kernels with a unit s
    partner) against the host engine's: the fitted tensor of (H2O)_16 (naux 1856, s..f
    auxiliary shells) agrees to 1e-11 of its largest element."""
    mol = M.water_cluster((4, 2, 2), seed=4)
    announce("metric (P|Q) (H2O)_16 cc-pVDZ-jkfit: device vs host")
    info_d, info_h = {}, {}
    Bd, _, td = _tensor(mol, "cc-pvdz", True, key="PSI4AMD_HOST_METRIC", info=info_d)
    Bh, _, th = _tensor(mol, "cc-pvdz", False, key="PSI4AMD_HOST_METRIC", info=info_h)
    e = _rel(Bd, Bh)
    print(f"  |B_dev_metric - B_host_metric|/max {e:.2e}; metric phase {info_d['setup_ms']['metric']:.1f} ms "
          f"(device) vs {info_h['setup_ms']['metric']:.1f} ms (host)", flush=True)
    assert e < 1e-11


def _schwarz(mol, basis, device, puream=None):
    from psi4_amd.basisset import BasisSet
    from psi4_amd.core import JK
    prim = BasisSet.build(mol, basis, puream=puream)
    aux = BasisSet.build(mol, basis + "-jkfit", puream=puream)
    old = os.environ.pop("PSI4AMD_HOST_SCHWARZ", None)
    if not device:
        os.environ["PSI4AMD_HOST_SCHWARZ"] = "1"
    try:
        jk = JK.build(prim, aux)
        t0 = time.perf_counter()
        mask = jk.pair_mask()
        dt = time.perf_counter() - t0
        fun, shell, mx, on_dev = jk.schwarz_values()
        nsig = jk.info()["nsig_pairs"]
    finally:
        os.environ.pop("PSI4AMD_HOST_SCHWARZ", None)
        if old is not None:
            os.environ["PSI4AMD_HOST_SCHWARZ"] = old
    assert on_dev == device
    return mask, fun, shell, mx, nsig, dt


@pytest.mark.parametrize("case", ["water-tz", "water-tz-cart", "cluster8-dz", "ar5-dz", "cluster24-dz"])
def test_device_schwarz_matches_host(gpu_available, case):
    """The Schwarz diagonal |(mn|mn)| of the screening (dfhelper.cc:329-356) computed on the device
    equals the host engine's element by element (1e-12 relative, every class s..f, solid-harmonic
    and Cartesian), and the screening built from it -- function-pair mask, pair count -- is the
    host's bit for bit; Ar5 keeps psi4's 1962 significant pairs (tests/pytests/test_jkmemory.py:37-45)."""
    if case.startswith("water-tz"):
        mol, basis = M.water_zmat(0.96, 104.5), "cc-pvtz"
    elif case == "ar5-dz":
        mol, basis = M.Molecule.from_string("Ar 0 0 0\nAr 0 0 5\nAr 0 0 15\nAr 0 0 25\nAr 0 0 35"), "cc-pvdz"
    elif case == "cluster8-dz":
        mol, basis = M.water_cluster((2, 2, 2), seed=9), "cc-pvdz"
    else:
        mol, basis = M.water_cluster((3, 4, 2), seed=0), "cc-pvdz"
    puream = False if case.endswith("cart") else None
    announce(f"Schwarz diagonal {case}: device")
    md, fd, sd, xd, nd, td = _schwarz(mol, basis, True, puream)
    announce(f"Schwarz diagonal {case}: host")
    mh, fh, sh, xh, nh, th = _schwarz(mol, basis, False, puream)
    err = np.abs(fd - fh)
    # relative error where it decides the mask (values above a tenth of the threshold
    # cutoff^2 / max), and the closest any value comes to the threshold
    tol = 1e-24 / xh
    big = fh >= 0.1 * tol
    rel = float((err[big] / fh[big]).max())
    margin = float(np.abs(fh[big] / tol - 1.0).min())
    print(f"  {case}: nbf {fd.shape[0]}, max |dev-host|/host {rel:.2e} over values >= tol/10, closest value "
          f"to tol {margin:.1e} relative, max {xd:.6e}; nsig {nd} (host {nh}); "
          f"screening {td * 1e3:.1f} ms device, {th * 1e3:.1f} ms host", flush=True)
    assert np.all(err <= 1e-12 * fh + 1e-16 * xh)
    assert abs(xd - xh) <= 1e-13 * xh
    assert np.all(np.abs(sd - sh) <= 1e-12 * sh + 1e-16 * xh)
    assert np.array_equal(md, mh) and nd == nh
    if case == "ar5-dz":
        assert nd == 1962


def test_setup_on_a_torch_stream_matches_own_stream(gpu_available):
    """Set-up (Schwarz diagonal, metric, its factor, three-index integrals, fit) on a torch
    stream, which does not synchronise with the legacy stream, gives the B of the handle's own
    stream: every copy between device and pageable host memory waits for the work queued
    before it and is complete when the call returns."""
    import torch
    from psi4_amd.basisset import BasisSet
    from psi4_amd.core import JK
    mol = M.water_cluster((2, 2, 1), seed=5)
    prim = BasisSet.build(mol, "cc-pvdz")
    aux = BasisSet.build(mol, "cc-pvdz-jkfit")

    def build(stream):
        jk = JK.build(prim, aux, memory=8_000_000_000)
        jk.set_print(0)
        if stream is not None:
            jk.set_stream(stream.cuda_stream)
        jk.initialize()
        torch.cuda.synchronize()
        return jk.get_B_dense(0, jk.info()["naux"])

    announce("set-up on a torch stream vs the handle's own stream, (H2O)_4 cc-pVDZ")
    B_own = build(None)
    B_torch = build(torch.cuda.Stream())
    assert np.abs(B_own).max() > 0
    assert _rel(B_torch, B_own) < 1e-14
