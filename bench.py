#!/usr/bin/env python3
"""Headline benchmark of the DF-JK path on one MI355X.

Metric: J+K builds/s at nbf=1500, naux=6000, nocc=300 (mixed vs FP64), |dE| vs FP64.
Workload: the seeded synthetic tensor-level case -- B_Qmn = s_Q exp(-|m-n|/64)
(G_Qmn+G_Qnm)/2 generated on the device, C_occ = first 300 columns of the QR
factor of a seeded N(0,1) 1500x1500 matrix.

A mixed-precision step is the Fock build the way SCF uses it: J(D_new), K(D_new)
from FP64 J, K of the previous density plus J(dD), K(dD) computed in the mixed
precision, with dD = C_new C_new^T - C_old C_old^T passed in the cross form
C_left = [A, V], C_right = [V, A] (A = (C_new + C_old)/2, V = C_new - C_old).
The FP64 step is one plain symmetric J+K build of D_new.  C is resident in HBM
and J, K are left in HBM.  Full builds in FP32 and 3xBF16 are timed beside them.

Printed: one JSON line (rank 0) with the headline precision's step rate as
``value``, the FP64 builds/s beside it, |dE| = |E_step - E_fp64| with
E = tr(D(2J-K)), the roofline of the step's dominant kernel (HIP events on the
library's stream over the timed region), and the CPU oracle timed on a
bounded Q-slice of the same workload (J and K are additive over Q).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X dense peaks in algorithmic TFLOP/s: FP64 MFMA, FP32 MFMA, and the BF16 MFMA
# peak divided by the three products of the 3-term split.
PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0 / 3}
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="bf16x3", help="headline (mixed) precision: fp32|bf16x3|fp64")
    ap.add_argument("--no-incremental", action="store_true",
                    help="time full builds only (the headline is then the full mixed build)")
    ap.add_argument("--incremental-only", action="store_true",
                    help="time only the incremental mixed step (no full builds, no |dE|): for counter runs")
    ap.add_argument("--cpu-sample", type=int, default=1800, help="aux functions in the CPU oracle slice (0: skip)")
    ap.add_argument("--nbf", type=int, default=None)
    ap.add_argument("--naux", type=int, default=None)
    ap.add_argument("--nocc", type=int, default=None)
    ap.add_argument("--skip-fp64", action="store_true", help="do not time the FP64 reference build")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run a device copy of this many GiB (HBM counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--scf-cluster", type=int, nargs=3, default=None, metavar=("NX", "NY", "NZ"),
                    help="instead of the synthetic build: RHF on a seeded water cluster of NX*NY*NZ molecules, "
                         "plain FP64 vs incremental mixed precision, then the energy adjoint")
    ap.add_argument("--cpu-check", action="store_true",
                    help="with --scf-cluster: the CPU oracle RHF (nbf <= 400), or the oracle J/K and energy at the "
                         "converged orbitals on the device's tensor (larger clusters)")
    return ap.parse_args()


def scf_workload(args):
    """RHF on a seeded water cluster (cc-pVDZ / cc-pVDZ-jkfit): plain FP64 GPU JK vs the incremental
    mixed-precision builds with FP64 finishing iterations, then dE/dC and dE/dB on the device."""
    import numpy as np
    import torch
    from psi4_amd.molecule import water_cluster
    from psi4_amd.scf import RHF
    nx, ny, nz = args.scf_cluster
    mol = water_cluster((nx, ny, nz), seed=0)
    prec = args.precision if args.precision != "fp64" else "bf16x3"
    log(f"(H2O)_{nx * ny * nz}: setup (integrals, device fit)")
    t0 = time.perf_counter()
    plain = RHF(mol)
    jk = plain.jk()
    t_setup = time.perf_counter() - t0
    info = jk.info()
    log(f"nbf {info['nbf']} naux {info['naux']}, setup {t_setup:.1f} s; warm-up; plain FP64 RHF")
    # untimed warm-up, so neither timed run pays one-time costs the other reuses: the JK
    # workspaces of a build at the occupied width, the device eigensolver and GEMM set-up
    nocc_w = mol.nelectron() // 2
    Cw = np.linalg.qr(np.random.default_rng(0).standard_normal((info["nbf"], nocc_w)))[0]
    jk.C_clear()
    jk.C_left_add(Cw)
    jk.compute()
    Aw = torch.rand((info["nbf"], info["nbf"]), dtype=torch.float64, device="cuda")
    torch.linalg.eigh(Aw + Aw.T)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rp = plain.compute_energy()
    t_plain = time.perf_counter() - t0
    log(f"plain: E {rp.energy:.10f} in {rp.iterations} iterations, {t_plain:.2f} s")
    mixed = RHF(mol, jk=jk, mixed_precision=prec)
    t0 = time.perf_counter()
    rm = mixed.compute_energy()
    t_mixed = time.perf_counter() - t0
    kinds = [h["build"] for h in rm.history]
    log(f"mixed ({prec}): E {rm.energy:.10f} in {rm.iterations} iterations ({kinds.count(prec)} mixed builds), "
        f"{t_mixed:.2f} s")
    # adjoint on the device: E, dE/dC, dE/dB; check <B, dE/dB> = 2E and <C, dE/dC> = 4E
    C = torch.from_numpy(np.ascontiguousarray(rm.C[:, :mixed.nocc])).cuda()
    nbf, naux = info["nbf"], info["naux"]
    dC = torch.empty_like(C)
    dB = torch.empty((naux, nbf, nbf), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E2 = jk.energy_adjoint_device(C.data_ptr(), C.shape[1], dC.data_ptr(), dB.data_ptr())
    t_adj = time.perf_counter() - t0
    Bd = torch.from_numpy(jk.get_B_dense(0, naux)).cuda()
    euler_B = float((Bd * dB).sum()) / (2 * E2) - 1.0
    euler_C = float((C * dC).sum()) / (4 * E2) - 1.0
    del Bd, dB
    out = {"workload": f"RHF (H2O)_{nx * ny * nz} cc-pVDZ/cc-pVDZ-jkfit, seeded cluster",
           "nbf": nbf, "naux": naux, "nocc": int(mixed.nocc), "setup_s": round(t_setup, 2),
           "setup_ms_by_phase": {k: round(v, 1) for k, v in info["setup_ms"].items()},
           "plain_fp64": {"E": rp.energy, "iterations": rp.iterations, "wall_s": round(t_plain, 3),
                          "jk_s": round(rp.timings["jk_s"], 3),
                          "jk_ms_per_iteration": [round(h["jk_s"] * 1e3, 2) for h in rp.history]},
           "mixed": {"precision": prec, "E": rm.energy, "iterations": rm.iterations,
                     "mixed_builds": kinds.count(prec), "wall_s": round(t_mixed, 3),
                     "jk_s": round(rm.timings["jk_s"], 3),
                     "jk_ms_per_iteration": [[h["build"], round(h["jk_s"] * 1e3, 2)] for h in rm.history]},
           "dE_mixed_vs_plain": abs(rm.energy - rp.energy),
           "adjoint": {"wall_s": round(t_adj, 3), "E_2": E2, "rel_<B,dE/dB>/2E-1": euler_B,
                       "rel_<C,dE/dC>/4E-1": euler_C}}
    if args.cpu_check and nbf > 400:
        # large clusters: the oracle's J/K contractions (compute_J_symm, compute_K) on the
        # device's fitted tensor, Q-slice by Q-slice on the host cores, at each run's
        # converged orbitals; E_cpu = E_nuc + 2<D,H> + <D, 2J - K>
        from oracle.dfhelper import compute_J_symm, compute_K, prepare_sparsity
        from psi4_amd import ints
        H = ints.kinetic(plain.basis) + ints.potential(plain.basis, mol)
        enuc = mol.nuclear_repulsion_energy()

        def cpu_energy(Cfull, nocc):
            Co = np.ascontiguousarray(Cfull[:, :nocc])
            Dh = Co @ Co.T
            Jc = np.zeros((nbf, nbf))
            Kc = np.zeros((nbf, nbf))
            for qa in range(0, naux, 200):
                qb = min(naux, qa + 200)
                Bq = jk.get_B_dense(qa, qb)
                sp = prepare_sparsity(np.ones((nbf, nbf)), np.ones((1, 1)), 1.0, 1.0, qb - qa)
                Ppq = np.ascontiguousarray(Bq.transpose(1, 0, 2)).reshape(-1)
                del Bq
                compute_J_symm(Ppq, sp, Dh, Jc)
                compute_K(Ppq, sp, Co, Co, Kc, True)
                if qa % 1000 == 0:
                    log(f"  CPU oracle J/K: aux functions {qa}..{qb} of {naux}")
            return enuc + 2.0 * np.vdot(Dh, H) + np.vdot(Dh, 2.0 * Jc - Kc)
        log("CPU oracle J/K at the converged orbitals (device B, host contractions)")
        t0 = time.perf_counter()
        Ep = cpu_energy(rp.C, plain.nocc)
        Em = cpu_energy(rm.C, mixed.nocc)
        out["cpu_oracle"] = {"kind": "oracle compute_J_symm + compute_K on the device's fitted B at the converged C",
                             "E_at_plain_C": Ep, "E_at_mixed_C": Em, "wall_s": round(time.perf_counter() - t0, 1),
                             "dE_plain_vs_cpu": abs(rp.energy - Ep), "dE_mixed_vs_cpu": abs(rm.energy - Em)}
    elif args.cpu_check:
        from oracle.dfhelper import MemDFJKOracle
        from oracle.rhf import rhf
        from psi4_amd import ints
        log("CPU oracle RHF (MemDFJK restatement, Cholesky fit)")
        t0 = time.perf_counter()
        ora = MemDFJKOracle(plain.basis, plain.aux, fit="cholesky")
        S = ints.overlap(plain.basis)
        H = ints.kinetic(plain.basis) + ints.potential(plain.basis, mol)

        def jkf(Cx):
            _, Jo, Ko = ora.compute([Cx])
            return Jo[0], Ko[0]
        ro = rhf(H, S, mol.nelectron() // 2, mol.nuclear_repulsion_energy(), jkf)
        out["cpu_oracle"] = {"E": ro["E"], "wall_s": round(time.perf_counter() - t0, 1),
                             "dE_mixed_vs_cpu": abs(rm.energy - ro["E"]), "dE_plain_vs_cpu": abs(rp.energy - ro["E"])}
    print(json.dumps(out), flush=True)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "only one GPU is measured: J and K are additive over the auxiliary index, so "
                                   "several GPUs would each hold a Q-slice of B and sum J/K, which is not "
                                   "built or measured here", "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    if args.scf_cluster:
        return scf_workload(args)

    if args.copy_probe > 0:
        n = int(args.copy_probe * (1 << 30)) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read_per_launch": n * 8, "bytes_written_per_launch": n * 8,
                          "launches": 3}), flush=True)
        return

    from psi4_amd.core import JK, PREC_BF16X3, PREC_FP32, PREC_FP64
    from psi4_amd.workloads import SYNTHETIC, jk_flops, synthetic_C

    nbf = args.nbf or SYNTHETIC["nbf"]
    naux = args.naux or SYNTHETIC["naux"]
    nocc = args.nocc or SYNTHETIC["nocc"]
    dev = torch.device("cuda:0")
    prec_id = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3}
    headline = args.precision
    if headline not in prec_id:
        raise SystemExit(f"unknown precision {headline}")

    log(f"synthetic workload nbf={nbf} naux={naux} nocc={nocc}: generating B on the device")
    jk = JK.build_tensor(nbf, naux)
    jk.set_print(0)
    # one stream for torch's work and the library's kernels (default priority)
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    jk.set_stream(stream.cuda_stream)
    jk.initialize()
    jk.fill_synthetic(SYNTHETIC["seed_B"])
    C_h = synthetic_C(nbf, nocc, SYNTHETIC["seed_C"])
    C = torch.from_numpy(C_h).to(dev)
    D = torch.empty(nbf, nbf, dtype=torch.float64, device=dev)
    J = torch.empty_like(D)
    K = torch.empty_like(D)
    torch.cuda.synchronize()

    def build():
        jk.compute_device([C.data_ptr()], [], [nocc], [D.data_ptr()], [J.data_ptr()], [K.data_ptr()], False)

    def energy():
        return float(torch.sum(D * (2.0 * J - K)).item())

    results = {}
    order = [headline] + [p for p in ("fp32", "fp64") if p != headline and not (p == "fp64" and args.skip_fp64)]
    if args.incremental_only:
        order = []
        D.copy_(C @ C.T)
    for p in order:
        jk.set_precision(prec_id[p])
        jk.set_profiling(True)
        log(f"{p}: {args.warmup} warmup + {args.steps} timed builds")
        for _ in range(args.warmup):
            build()
        torch.cuda.synchronize()
        tsum = {"J": 0.0, "HT": 0.0, "KGEMM": 0.0, "other": 0.0, "n_ht": 0, "n_kgemm": 0}
        t0 = time.perf_counter()
        for _ in range(args.steps):
            build()
            t = jk.timings()
            for k in tsum:
                tsum[k] += t[k]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        results[p] = dict(s_per_build=dt / args.steps, E=energy(), timings=tsum)
        log(f"{p}: {dt / args.steps * 1e3:.2f} ms/build, E={results[p]['E']:.12f}, kernels(ms)="
            f"J {tsum['J'] / args.steps:.2f} HT {tsum['HT'] / args.steps:.2f} KGEMM {tsum['KGEMM'] / args.steps:.2f}")
    jk.set_profiling(False)

    # Incremental build (the way the mixed path is used inside SCF): J(dD), K(dD) added to
    # FP64 J/K of D_old, in the cross form dD = A V^T + V A^T with A = (C_new + C_old)/2,
    # V = C_new - C_old after aligning C_old to C_new (orthogonal Procrustes), i.e.
    # C_left=[A, V], C_right=[V, A]: every low-precision product carries the small factor V,
    # and the engine contracts it as a symmetric K.  The same build in FP64 checks the
    # identity K(D_old) + K(dD) = K(D_new) at this size.
    inc = None
    tinc = None
    if not args.no_incremental and headline != "fp64":
        rng = np.random.default_rng(SYNTHETIC["seed_C"] + 1)
        Cold_h = np.linalg.qr(C_h + 1e-4 * rng.standard_normal(C_h.shape))[0]
        u, _, vt = np.linalg.svd(Cold_h.T @ C_h)
        Cold_h = Cold_h @ (u @ vt)
        V_h = C_h - Cold_h
        A_h = 0.5 * (C_h + Cold_h)
        Cold = torch.from_numpy(np.ascontiguousarray(Cold_h)).to(dev)
        Jold = torch.empty_like(D)
        Kold = torch.empty_like(D)
        jk.set_precision(PREC_FP64)
        jk.compute_device([Cold.data_ptr()], [], [nocc], [], [Jold.data_ptr()], [Kold.data_ptr()], False)
        Cl = torch.from_numpy(np.ascontiguousarray(np.hstack([A_h, V_h]))).to(dev)
        Cr = torch.from_numpy(np.ascontiguousarray(np.hstack([V_h, A_h]))).to(dev)
        Ji = torch.empty_like(D)
        Ki = torch.empty_like(D)

        # the SCF driver declares the cross form, so the engine skips its column test (no host wait)
        jk.set_pair_form("CROSS")

        def inc_build():
            Ji.copy_(Jold)
            Ki.copy_(Kold)
            jk.compute_device([Cl.data_ptr()], [Cr.data_ptr()], [2 * nocc], [], [Ji.data_ptr()], [Ki.data_ptr()], True)

        E_inc = {}
        if "fp64" in results:
            jk.set_precision(PREC_FP64)
            inc_build()
            E_inc["fp64"] = float(torch.sum(D * (2.0 * Ji - Ki)).item())
        jk.set_precision(prec_id[headline])
        log(f"incremental {headline}: {args.warmup} warmup + {args.steps} timed dD builds (ncol={2 * nocc})")
        jk.set_profiling(True)
        for _ in range(args.warmup):
            inc_build()
        torch.cuda.synchronize()
        tinc = {"J": 0.0, "HT": 0.0, "KGEMM": 0.0, "other": 0.0, "n_ht": 0, "n_kgemm": 0}
        t0 = time.perf_counter()
        for _ in range(args.steps):
            inc_build()
            t = jk.timings()
            for k in tinc:
                tinc[k] += t[k]
        torch.cuda.synchronize()
        dt_inc = (time.perf_counter() - t0) / args.steps
        jk.set_profiling(False)
        E_inc[headline] = float(torch.sum(D * (2.0 * Ji - Ki)).item())
        dD = float(np.abs(C_h @ C_h.T - Cold_h @ Cold_h.T).max())
        ef = results["fp64"]["E"] if "fp64" in results else None
        inc = {"builds_per_s": round(1.0 / dt_inc, 4), "ms_per_build": round(dt_inc * 1e3, 3),
               "speedup_vs_fp64_full": (round(results["fp64"]["s_per_build"] / dt_inc, 3) if ef is not None else None),
               "precision": headline, "form": "C_left=[A,V], C_right=[V,A]; A=(C_new+C_old)/2, V=C_new-C_old",
               "max_abs_dD": dD, "max_abs_V": float(np.abs(V_h).max()),
               "dE_vs_fp64": (abs(E_inc[headline] - ef) if ef is not None else None),
               "dE_fp64_incremental_vs_fp64_full": (abs(E_inc["fp64"] - ef) if ef is not None else None)}
        inc["kernel_ms_per_build"] = {k: round(v / args.steps, 4) for k, v in tinc.items() if not k.startswith("n_")}
        log(f"incremental {headline}: {dt_inc * 1e3:.2f} ms/build, |dD|max={dD:.2e}, dE={inc['dE_vs_fp64']}, "
            f"FP64 incremental dE={inc['dE_fp64_incremental_vs_fp64_full']}, kernels(ms)=J "
            f"{tinc['J'] / args.steps:.2f} HT {tinc['HT'] / args.steps:.2f} KGEMM {tinc['KGEMM'] / args.steps:.2f}")
        jk.set_precision(prec_id[headline])

    # the headline step: the incremental mixed build when it ran, else the full build
    step_kind = "incremental" if inc is not None else "full"
    if step_kind == "incremental":
        s = dt_inc
        tt = tinc
        dE_step = inc["dE_vs_fp64"]
        ncol_ht = 2 * nocc      # C_left = [A, V]; the contraction forms X_A X_V^T (nocc columns) over every tile
    else:
        hr = results[headline]
        s = hr["s_per_build"]
        tt = hr["timings"]
        dE_step = abs(hr["E"] - results["fp64"]["E"]) if "fp64" in results else None
        ncol_ht = nocc

    # roofline of the dominant kernel class of the headline step
    fl_j, fl_ht, fl_kc = jk_flops(nbf, naux, ncol_ht, symmetric=True)
    dom = "HT" if tt["HT"] >= tt["KGEMM"] else "KGEMM"
    nlaunch = tt["n_ht"] if dom == "HT" else tt["n_kgemm"]
    flops_per_launch = (fl_ht if dom == "HT" else fl_kc) * args.steps / max(1, nlaunch)
    ms_per_launch = tt[dom] / max(1, nlaunch)
    achieved = flops_per_launch / (ms_per_launch * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK[headline], "unit": "TFLOP/s",
                "frac": round(achieved / PEAK[headline], 4), "traffic": None,
                "kernel": ({"fp64": "k_ht_mfma<double>", "fp32": "k_ht_mfma<float>", "bf16x3": "k_ht_bf3"}[headline]
                           + " (half-transform X_Q = B_Q C)") if dom == "HT" else
                          ({"fp64": "k_kgemm_mfma<double>", "fp32": "k_kgemm_mfma<float>", "bf16x3": "k_kgemm_bf3"}[headline]
                           + " (K = sum_Q X_Q X_Q^T)"),
                "launches_per_build": nlaunch / args.steps, "ms_per_launch": round(ms_per_launch, 4),
                "algorithmic_flops_per_launch": flops_per_launch}
    # J path: bandwidth view (two streaming passes over packed B: FP64, or its FP32 copy in
    # the mixed precisions)
    # algorithmic bytes 2 * naux * n_sym * s_B (n_sym = nbf(nbf+1)/2); the kernels stream the
    # stored 32x32 tiles, padded_factor more
    ntp = jk.info()["ntile_pairs"]
    s_b = 8 if headline == "fp64" else 4
    j_bytes = 2 * naux * (nbf * (nbf + 1) // 2) * s_b
    j_ms = tt["J"] / args.steps
    j_roof = {"bound": "hbm", "achieved_GBs": round(j_bytes / (j_ms * 1e-3) / 1e9, 1), "peak_GBs": HBM_PEAK_GBS,
              "bytes_per_build": j_bytes, "ms_per_build": round(j_ms, 4),
              "padded_factor": round(2 * naux * ntp * 1024 * s_b / j_bytes, 4)}

    # measured HBM traffic of that kernel (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, calibrated
    # against a copy kernel), recorded per dispatch in profiles/ for this workload
    tfile = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                         f"traffic_synthetic_nbf{nbf}_naux{naux}_nocc{nocc}.json")
    kname = roofline["kernel"].split(" ")[0]
    if os.path.exists(tfile):
        rec = json.load(open(tfile)).get("steps", {}).get(step_kind, {}).get(kname)
        if rec:
            # bytes read + written by all dispatches of one build, over the aux functions they cover
            per_q = (rec["read_bytes"] + rec["write_bytes"]) / rec["q_covered"]
            q_per_launch = naux * args.steps / max(1, nlaunch)
            roofline["traffic"] = round(per_q * q_per_launch)
            roofline["traffic_source"] = f"{os.path.basename(tfile)} ({rec.get('profiled', 'rocprofv3 --pmc')})"

    # CPU oracle on a bounded Q-slice of the same workload (J and K are additive over Q:
    # the slice is processed in chunks of 200 aux functions to bound host memory)
    cpu = None
    if args.cpu_sample > 0:
        qs = min(args.cpu_sample, naux)
        log(f"CPU oracle on a Q-slice of {qs} aux functions")
        from oracle.dfhelper import compute_J_symm, compute_K, prepare_sparsity
        Dh = C_h @ C_h.T
        Jc = np.zeros((nbf, nbf))
        Kc = np.zeros((nbf, nbf))
        tcpu = 0.0
        for qa in range(0, qs, 200):
            qb = min(qs, qa + 200)
            Bq = jk.get_B_dense(qa, qb)
            sp = prepare_sparsity(np.ones((nbf, nbf)), np.ones((1, 1)), 1.0, 1.0, qb - qa)
            Ppq = np.ascontiguousarray(Bq.transpose(1, 0, 2)).reshape(-1)
            del Bq
            t0 = time.perf_counter()
            compute_J_symm(Ppq, sp, Dh, Jc)
            compute_K(Ppq, sp, C_h, C_h, Kc, True)
            tcpu += time.perf_counter() - t0
            del Ppq
        log(f"CPU oracle: {tcpu:.2f} s of J/K work on {qs} aux functions")
        try:
            from threadpoolctl import threadpool_info
            cores = max([i.get("num_threads", 1) for i in threadpool_info()] + [1])
        except Exception:
            cores = os.cpu_count()
        # GPU on the same slice (FP64) for a same-run parity number
        jk.set_precision(PREC_FP64)
        jk.set_aux_range(0, qs)
        build()
        torch.cuda.synchronize()
        dj = float(np.abs(J.cpu().numpy() - Jc).max())
        dk = float(np.abs(K.cpu().numpy() - Kc).max())
        jk.set_aux_range(0, naux)
        cpu = {"value": round(1.0 / (tcpu * naux / qs), 5), "unit": "builds/s", "cores": cores, "kind": "port",
               "sample": f"oracle compute_J_symm+compute_K on Q-slice [0,{qs}) of {naux}, {tcpu:.2f} s, "
                         f"extrapolated x{naux / qs:.1f} (J, K additive over Q); host {platform.processor() or 'x86_64'}",
               "slice_maxabs_dJ": dj, "slice_maxabs_dK": dk}
        log(f"CPU slice {tcpu:.2f}s -> {cpu['value']} builds/s; GPU-vs-CPU slice |dJ|={dj:.2e} |dK|={dk:.2e}")

    out = {
        "metric": "J+K builds/s at nbf=1500,naux=6000,nocc=300 (mixed vs FP64); |dE| vs FP64",
        "value": round(1.0 / s, 4),
        "unit": "builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32 (K), f64 (J, accumulation)",
                  "bf16x3": "bf16x3 split, f32 accumulate (K); f64 (J, accumulation)"}[headline],
        "data": "synthetic (device Philox4x32-10 B, seeded QR C_occ)",
        "config": {"workload": (f"synthetic tensor-level DF-JK nbf={nbf} naux={naux} nocc={nocc}, "
                                + ("incremental Fock build: FP64 J/K of D_old + J/K(dD) in " + headline
                                   + " (cross form, 2*nocc columns)" if step_kind == "incremental"
                                   else f"full {headline} J+K build")),
                   "step": step_kind, "nbf": nbf, "naux": naux, "nocc": nocc, "precision": headline,
                   "parallelism": "single GPU"},
        "fp64_builds_per_s": round(1.0 / results["fp64"]["s_per_build"], 4) if "fp64" in results else None,
        "speedup_vs_fp64": (round(results["fp64"]["s_per_build"] / s, 3) if "fp64" in results else None),
        "dE_vs_fp64": dE_step,
        "E_fp64": results.get("fp64", {}).get("E"),
        "by_precision": {p: {"builds_per_s": round(1.0 / r["s_per_build"], 4), "E": r["E"],
                             "dE_vs_fp64": (abs(r["E"] - results["fp64"]["E"]) if "fp64" in results else None),
                             "kernel_ms_per_build": {k: round(v / args.steps, 4) for k, v in r["timings"].items()
                                                     if not k.startswith("n_")}}
                         for p, r in results.items()},
        "incremental": inc,
        "kernel_ms_per_build": {k: round(v / args.steps, 4) for k, v in tt.items() if not k.startswith("n_")},
        "roofline": roofline,
        "roofline_J": j_roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
