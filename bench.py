"""Frequency-point solves/s of the dense extraction on one MI355X.

Workload: the reference's scaling fixture meander(49) at 2 x 2 filaments per
segment (fasterhenry/benches/support.rs:104-127, benches/scaling.rs:91):
9 800 filaments, 7 350 internal loops, 1 port, swept over 41 log-spaced
frequencies from 10 kHz to 1 GHz.  A step is one sweep of the 41
frequencies over the assembled system resident in HBM (forming
R + jwL, factorizing, Schur complement, Z back to the host); assembly (fill +
congruence) is timed once and reported separately.

Prints one JSON line (the driver's contract).  --gpus N > 1 exits: this
build measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAKS_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(system, omega_hz: float, threads: int) -> dict:
    """The oracle (the CPU FP64 restatement of dense.rs lu_solve + the
    Schur complement of solve.rs:1154-1164) on one frequency point of the
    same mesh blocks, all host cores."""
    import oracle

    b = system.blocks()
    omega = 2.0 * np.pi * omega_hz
    ee = b["r_ee"] + 1j * omega * b["l_ee"]
    ep = b["r_ep"] + 1j * omega * b["l_ep"]
    pp = b["r_pp"] + 1j * omega * b["l_pp"]
    del b
    t0 = time.perf_counter()
    x = oracle.lu_solve(ee, ep, threads=threads)
    z = pp - ep.T @ x
    dt = time.perf_counter() - t0
    return {"seconds": dt, "z": z}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--bars", type=int, default=49)
    ap.add_argument("--freqs", type=int, default=41)
    ap.add_argument("--precision", default="fp32_refined", choices=["fp64", "fp32_refined"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the plain-FP64 sweep the refined result is compared with")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; run with --gpus 1",
                          "n_gpus": args.gpus}))
        return 2

    from fasterhenry_amd import MeshSystem, fixtures, model, native

    g, ports, d = fixtures.meander_problem(args.bars)
    packed = model.pack(g, ports, d)
    freqs = fixtures.log_frequencies(1e4, 1e9, args.freqs)
    print(f"# assembling meander({args.bars}): {packed.filament_count} filaments", flush=True)
    t0 = time.perf_counter()
    system = MeshSystem.assemble_packed(packed, ports, args.precision)
    assembly_s = time.perf_counter() - t0
    counts = system.counts()
    print(f"# assembled in {assembly_s:.2f} s: m={counts.internal_meshes} p={counts.ports}",
          flush=True)

    for w in range(args.warmup):
        system.sweep_raw(freqs)
        print(f"# warmup {w + 1}/{args.warmup} done", flush=True)
    system.set_profiling(True)
    infos = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        z, info = system.sweep_raw(freqs)
        infos.append(info.as_dict())
        print(f"# step {s + 1}/{args.steps}: {info.solve_s:.3f} s", flush=True)
    elapsed = time.perf_counter() - t0
    system.set_profiling(False)
    value = args.freqs * args.steps / elapsed

    upd_ms = sum(i["update_kernel_ms"] for i in infos)
    upd_n = sum(i["update_kernel_launches"] for i in infos)
    upd_flops = sum(i["update_kernel_flops"] for i in infos)
    upd_unit_flops = sum(i["update_kernel_unit_flops"] for i in infos)
    unit_kind = "fp64" if args.precision == "fp64" else "fp32"
    # The roofline is priced on the units the dominant kernel runs on: FP64 or
    # FP32 matrix cores, or the bf16 matrix cores for split-bf16 updates (three
    # bf16 products per FP32-equivalent product).
    arith = infos[-1]["factor_arith"]
    peak_kind = {0: "fp64", 1: "fp32", 2: "bf16"}[arith]
    peak = PEAKS_TFLOPS[peak_kind]
    achieved = upd_unit_flops / (upd_ms * 1e-3) / 1e12 if upd_ms > 0 else 0.0
    achieved_equiv = upd_flops / (upd_ms * 1e-3) / 1e12 if upd_ms > 0 else 0.0

    extra = {}
    m = counts.internal_meshes
    if args.precision != "fp64" and not args.no_compare:
        system.set_precision("fp64")
        t1 = time.perf_counter()
        z64, _ = system.sweep_raw(freqs)
        t64 = time.perf_counter() - t1
        system.set_precision(args.precision)
        err = max(np.linalg.norm(z[k] - z64[k]) / np.linalg.norm(z64[k]) for k in range(len(freqs)))
        extra = {"z_rel_err_vs_fp64": err, "fp64_solves_per_s": args.freqs / t64,
                 "speedup_vs_fp64": (args.freqs / t64) and value / (args.freqs / t64)}

    cpu = None
    if not args.no_cpu:
        threads = os.cpu_count() or 1
        print(f"# cpu baseline: one frequency point, {threads} threads", flush=True)
        res = cpu_baseline(system, float(freqs[len(freqs) // 2]), threads)
        zg = z[len(freqs) // 2]
        cpu = {"value": 1.0 / res["seconds"], "unit": "freq-point solves/s", "cores": threads,
               "kind": "port",
               "sample": f"one frequency point ({freqs[len(freqs) // 2]:.4g} Hz) of the same "
                         f"m={m} system: the oracle's blocked complex LU (PANEL 32, partial "
                         f"pivoting) + Schur, OpenMP on {threads} threads of {cpu_model()}",
               "z_rel_err_gpu_vs_cpu": float(np.linalg.norm(zg - res["z"]) / np.linalg.norm(res["z"]))}

    out = {
        "metric": "freq-point solves/s @9.8k filaments: refined FP32 vs FP64 LU, Z err ≤1e-10",
        "value": value,
        "unit": "freq-point solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": unit_kind,
        "data": "synthetic: meander fixture of fasterhenry/benches/support.rs",
        "config": {"workload": f"meander({args.bars}) 2x2, {counts.filaments} filaments, "
                               f"{m} internal loops, {args.freqs} log freqs 10 kHz-1 GHz",
                   "precision": args.precision, "filaments": counts.filaments,
                   "internal_loops": m, "frequencies": args.freqs,
                   "assembly_s": assembly_s, "inductance_s": infos[0]["inductance_s"],
                   "factor_arith": {0: "fp64", 1: "fp32-mfma", 2: "fp32 split-bf16 mfma"}[
                       infos[-1]["factor_arith"]],
                   "refine_steps_max": infos[-1]["refine_steps_max"],
                   "refine_change_max": infos[-1]["refine_change_max"]},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "kernel": "outer trailing update of the L D L^T factorization "
                               f"({peak_kind} matrix cores)",
                     "achieved_equivalent_tflops": achieved_equiv,
                     "launches": upd_n, "avg_launch_ms": upd_ms / max(upd_n, 1)},
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(out))
    system.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
