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

# HBM bytes per update_left_kernel launch on the default workload (meander(49),
# 41 frequencies, refined, outer block 256), per frequency in the launch:
# the mean over a three-part sweep's 342 launches (41/3 frequencies each) of
# rocprofv3 FETCH_SIZE 590.5 MB + WRITE_SIZE 43.5 MB, each scaled by its
# reading for a plain 512 MiB device copy (bench.py --copy-mib 512: the copy
# kernel reads 262 155 KiB, factor 2.0; writes 524 288 KiB, factor 1.0).
# A launch of k frequencies moves k times this.
UPDATE_TRAFFIC_BYTES_PER_FREQUENCY = (590.5e6 + 43.5e6) / (41 / 3)

# HBM bytes per pFFT operator product (V = 2) on meander(157): the sum over
# the product's kernels (spread, line passes, column DFTs, gather, near
# field) of rocprofv3 FETCH_SIZE (3 593 MB) and WRITE_SIZE (1 382 MB) per
# product, over the dispatches of the 76 products only (not the set-up's
# line-pass timing), scaled as above.
PFFT_TRAFFIC_BYTES = 3593.2e6 + 1381.9e6


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


def cpu_end_to_end(threads: int) -> dict:
    """The reference's published CPU timing (docs/benchmarks.md:127-166):
    assemble + sweep at 1 MHz of the square meander at 2 x 2, here through
    the oracle at 960 and 2 400 filaments, on 1 thread and on `threads`."""
    import oracle
    from fasterhenry_amd import fixtures, model

    out = {"timed": "assemble + sweep(&[1 MHz]) of meander(bars) 2x2, oracle",
           "reference_1_thread_s": {"960": 0.84, "2400": 4.15},
           "reference_host": "AWS EC2 8 vCPU Xeon-class, 1 thread (different hardware)"}
    for bars in (15, 24):
        pk = model.pack(*fixtures.meander_problem(bars))
        row = {}
        for t in (1, threads):
            t0 = time.perf_counter()
            oracle.System(pk, t).sweep([1e6])
            row[f"{t}_thread_s" if t == 1 else f"{t}_threads_s"] = time.perf_counter() - t0
        out[str(pk.filament_count)] = row
        print(f"# cpu end to end {pk.filament_count} filaments: {row}", flush=True)
    return out


HBM_PEAK_GBS = 8000.0


def apply_bytes(st: dict, vectors: int, ncomp: int) -> float:
    """HBM bytes of one operator product on `vectors` vectors with the
    pruned FFT convolution: the transposed and direct projection CSR read
    once each (u32 index + f64 weight), the compact grid and the pruned
    intermediate arrays written and read once per line pass, the even
    eighth of the spectrum, the near-field CSR, x and y."""
    n, E, nnz = st["filaments"], st["projection_entries"], st["near_entries"]
    N0, N1, N2 = st["grid_dims"]
    M0, M1, M2 = st["fft_dims"]
    c = (vectors * ncomp + 1) // 2               # complex channels
    z = 8 if st["fft_kind"] == 1 else 16          # complex value
    G = N0 * N1 * N2
    a = c * N0 * N1 * M2 * z                      # after the axis-2 pass
    b = c * N0 * M1 * M2 * z                      # after the axis-1 pass
    spec = (M0 // 2 + 1) * (M1 // 2 + 1) * (M2 // 2 + 1) * (4 if st["fft_kind"] == 1 else 8)
    spread = 12 * E + 8 * G + c * G * z
    passes = (c * G * z + a) + (a + b) + (2 * b + spec) + (b + a) + (a + c * G * z)
    gather = 12 * E + 8 * n + c * G * z
    near = 12 * nnz + 8 * n + vectors * n * 8 * 3
    return float(spread + passes + gather + near)


def matrix_free(steps: int, warmup: int, dense_check=None) -> dict:
    """The matrix-free sweep: meander(157) at 2 x 2 (99 224 filaments,
    benches/scaling.rs:91), pFFT operator with the reference's default grid
    (auto, 8 cells per filament, order 5, radius 8) and an FP32 FFT
    convolution, FP64 near field and Arnoldi, GMRES restart 30, tolerance
    1e-8, 8 log frequencies 1 MHz - 1 GHz.  A step is one sweep (assembly
    timed once).  With dense_check = (packed, ports) of meander(49), the
    same sweep's iterative Z at 9 800 filaments is compared with the dense
    FP64 GPU solve."""
    from fasterhenry_amd import (IterativeSystem, MeshSystem, fixtures, iterative_params, model,
                                 pfft_params)

    freqs = fixtures.log_frequencies(1e6, 1e9, 8)
    params = iterative_params(pfft_params(fft_precision="fp32"), tolerance=1e-8, restart=30)
    g, ports, d = fixtures.meander_problem(157)
    packed = model.pack(g, ports, d)
    print(f"# matrix-free: assembling meander(157): {packed.filament_count} filaments", flush=True)
    t0 = time.perf_counter()
    it = IterativeSystem.assemble_packed(packed, ports, params)
    assembly_s = time.perf_counter() - t0
    st = it.operator_stats()
    counts = it.counts()
    print(f"# matrix-free: assembled in {assembly_s:.2f} s, grid {st['grid_dims']}, "
          f"padded {st['fft_dims']}", flush=True)
    for w in range(warmup):
        it.sweep_raw(freqs)
    it.set_profiling(True)
    infos = []
    t0 = time.perf_counter()
    for k in range(steps):
        z, info = it.sweep_raw(freqs)
        infos.append(info.as_dict())
        print(f"# matrix-free step {k + 1}/{steps}: {info.solve_s:.3f} s, "
              f"{info.gmres_iterations} GMRES iterations", flush=True)
    elapsed = time.perf_counter() - t0
    it.set_profiling(False)
    applies = sum(i["operator_applies"] for i in infos)
    apply_ms = sum(i["operator_apply_ms"] for i in infos) / max(applies, 1)
    ncomp = 2  # the meander's filaments run along x and y
    traffic = apply_bytes(st, 2, ncomp)
    achieved = traffic / (apply_ms * 1e-3) / 1e9 if apply_ms > 0 else 0.0
    out = {
        "metric": "matrix-free freq-point solves/s @99k filaments (pFFT + GMRES)",
        "value": len(freqs) * steps / elapsed,
        "unit": "freq-point solves/s",
        "ms_per_step": elapsed / steps * 1e3,
        "dtype": "fp32 FFT, fp64 near field and Arnoldi",
        "config": {"workload": f"meander(157) 2x2, {counts.filaments} filaments, "
                               f"{counts.internal_meshes} internal loops, 8 log freqs 1 MHz-1 GHz",
                   "assembly_s": assembly_s, "operator_build_s": infos[0]["inductance_s"],
                   "grid_dims": st["grid_dims"], "fft_dims": st["fft_dims"],
                   "projection_entries": st["projection_entries"],
                   "near_entries": st["near_entries"], "gmres_restart": 30,
                   "gmres_tolerance": 1e-8, "gmres_iterations": infos[-1]["gmres_iterations"],
                   "gmres_iterations_max": infos[-1]["gmres_iterations_max"],
                   "line_group": st["line_group"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": PFFT_TRAFFIC_BYTES,
                     "kernel": "pFFT operator product on the real and imaginary branch "
                               "currents (spread, 5 pruned FFT passes, gather, near field)",
                     "bytes_per_apply": traffic, "avg_apply_ms": apply_ms, "applies": applies},
    }
    if dense_check is not None:
        packed49, ports49 = dense_check
        it49 = IterativeSystem.assemble_packed(packed49, ports49, params)
        z49, _ = it49.sweep_raw(freqs)
        it49.close()
        dense = MeshSystem.assemble_packed(packed49, ports49, "fp64")
        zd, _ = dense.sweep_raw(freqs)
        dense.close()
        out["z_rel_err_vs_dense_9800"] = max(
            float(np.linalg.norm(z49[k] - zd[k]) / np.linalg.norm(zd[k])) for k in range(len(freqs)))
    it.close()
    return out


def batch(steps: int, warmup: int) -> dict:
    """The batched sweep: 48 seeded geometries (square spirals of 2-8 turns
    and busbar pairs, 384 - 3 825 filaments, surface-graded), each swept
    over 16 log frequencies 100 kHz - 3 GHz in the refined form, all through
    one fh_solve_batch call (assembly included: each problem is assembled,
    swept and released inside the call).  Also timed with one worker (the
    problems one after another) for the scheduling gain, and checked against
    plain FP64 solves of three problems."""
    from fasterhenry_amd import MeshSystem, fixtures, model, solve_batch

    problems = fixtures.batch_problems(48, 48)
    freqs = fixtures.log_frequencies(1e5, 3e9, 16)
    packed = [model.pack(*p) for p in problems]
    sizes = [pk.filament_count for pk in packed]
    print(f"# batch: {len(problems)} problems, {min(sizes)}-{max(sizes)} filaments "
          f"({sum(sizes)} total)", flush=True)
    for _ in range(warmup):
        solve_batch(problems, freqs)
    t0 = time.perf_counter()
    for k in range(steps):
        zs, infos = solve_batch(problems, freqs)
        print(f"# batch step {k + 1}/{steps} done", flush=True)
    elapsed = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    zs1, _ = solve_batch(problems, freqs, concurrency=1)
    sequential = time.perf_counter() - t0
    same = all(np.array_equal(a.view(np.float64), b.view(np.float64)) for a, b in zip(zs, zs1))
    err = 0.0
    for b in (0, 1, int(np.argmax(sizes))):
        g, ports, d = problems[b]
        system = MeshSystem.assemble_packed(packed[b], ports, "fp64")
        z64, _ = system.sweep_raw(freqs)
        system.close()
        err = max(err, max(float(np.linalg.norm(zs[b][k] - z64[k]) / np.linalg.norm(z64[k]))
                           for k in range(len(freqs))))
    m = [i["internal_meshes"] for i in infos]
    flops = sum(len(freqs) * (4.0 / 3.0) * float(x) ** 3 for x in m)
    return {
        "metric": "batched freq-point solves/s (48 uneven geometries x 16 freqs, assembly included)",
        "value": len(problems) * len(freqs) / elapsed,
        "unit": "freq-point solves/s",
        "ms_per_step": elapsed * 1e3,
        "dtype": "fp32 split-bf16 factorization + fp64 refinement",
        "config": {"workload": f"48 seeded spirals / busbar pairs, {min(sizes)}-{max(sizes)} "
                               f"filaments, 16 log freqs 100 kHz-3 GHz",
                   "filaments_total": sum(sizes), "precision": "fp32_refined",
                   "refine_steps_max": max(i["refine_steps_max"] for i in infos)},
        "sequential_ms": sequential * 1e3,
        "speedup_vs_sequential": sequential / elapsed,
        "identical_to_sequential": same,
        "z_rel_err_vs_fp64": err,
        "factorization_equivalent_tflops": flops / elapsed / 1e12,
    }


def adjoint(steps: int, warmup: int, fd_check: bool = True, cpu: bool = False) -> dict:
    """The gradient workload: a 6-turn square spiral (outer 2 mm, w 100 um,
    s 50 um, t 35 um), skin-depth graded at 1 GHz (0.5 delta, ratio 2) and
    frozen, 6 156 filaments; 12 parameters (per-turn width and spacing);
    dR/dp and dL/dp at 1, 10, 100, 500 MHz and 1 GHz through fh_grad (refined
    solve, filament currents, derivative fill contracted on the device).
    Checked against central differences of Z at relative step 1e-6 (Z from
    the same refined solve in its stationary form, 24 perturbed assemblies)."""
    from fasterhenry_amd import MeshSystem, fixtures

    freqs = [1e6, 1e7, 1e8, 5e8, 1e9]
    s = fixtures.AdjointSpiral()
    g, ports, d = s.problem()
    t0 = time.perf_counter()
    system = MeshSystem.assemble(g, ports, d, precision="fp32_refined")
    assembly_s = time.perf_counter() - t0
    counts = system.counts()
    d_nodes, d_widths = s.directions()
    print(f"# adjoint: {counts.filaments} filaments, {s.n_params} params", flush=True)
    for _ in range(warmup):
        system.grad(freqs, d_nodes, d_widths)
    t0 = time.perf_counter()
    for k in range(steps):
        z, dr, dl = system.grad(freqs, d_nodes, d_widths)
        print(f"# adjoint step {k + 1}/{steps} done", flush=True)
    elapsed = (time.perf_counter() - t0) / steps
    system.close()
    out = {
        "metric": "adjoint gradients/s (12 params x 5 freqs, dR/dp and dL/dp)",
        "value": 1.0 / elapsed,
        "unit": "gradients/s",
        "ms_per_step": elapsed * 1e3,
        "dtype": "fp64 derivative fill + fp32 split-bf16 factorization + fp64 refinement",
        "config": {"workload": f"6-turn square spiral 2 mm, w 100 um s 50 um t 35 um, skin-depth "
                               f"graded (1 GHz, 0.5 delta, ratio 2) frozen, {counts.filaments} "
                               f"filaments, 12 params, 5 freqs 1 MHz-1 GHz",
                   "filaments": counts.filaments, "internal_loops": counts.internal_meshes,
                   "params": s.n_params, "frequencies": len(freqs), "assembly_s": assembly_s},
    }
    if fd_check:
        w = 2 * np.pi * np.asarray(freqs)
        fr = np.zeros((len(freqs), s.n_params))
        fl = np.zeros_like(fr)
        t1 = time.perf_counter()
        for q in range(s.n_params):
            h = 1e-6 * s.nominal[q]
            zz = []
            for sgn in (1.0, -1.0):
                p = s.nominal.copy()
                p[q] += sgn * h
                sq = MeshSystem.assemble(*s.problem(p), precision="fp32_refined")
                zz.append(sq.grad(freqs, np.zeros((0,) + d_nodes.shape[1:]))[0][:, 0, 0])
                sq.close()
            dz = (zz[0] - zz[1]) / (2 * h)
            fr[:, q], fl[:, q] = dz.real, dz.imag / w
        fd_s = time.perf_counter() - t1
        er = np.abs(dr[:, :, 0, 0] - fr).max(axis=1) / np.abs(fr).max(axis=1)
        el = np.abs(dl[:, :, 0, 0] - fl).max(axis=1) / np.abs(fl).max(axis=1)
        out["fd_check"] = {"step": "central, relative 1e-6", "dR_rel_err_max": float(er.max()),
                           "dL_rel_err_max": float(el.max()),
                           "dR_rel_err_per_freq": [float(x) for x in er],
                           "dL_rel_err_per_freq": [float(x) for x in el],
                           "error_norm": "max over params of |adjoint - fd| / max over params |fd|, "
                                         "per frequency",
                           "fd_gradient_s_on_gpu": fd_s}
    if cpu:
        import oracle

        threads = os.cpu_count() or 1
        print(f"# adjoint cpu baseline: one oracle assembly + one frequency, {threads} threads",
              flush=True)
        t1 = time.perf_counter()
        osys = oracle.System(model_pack(s.problem()), threads)
        ta = time.perf_counter() - t1
        t1 = time.perf_counter()
        osys.sweep([freqs[2]])
        tf = time.perf_counter() - t1
        per_grad = (2 * s.n_params + 1) * (ta + len(freqs) * tf)
        out["cpu_baseline"] = {
            "value": 1.0 / per_grad, "unit": "gradients/s", "cores": threads, "kind": "port",
            "sample": f"one oracle assembly ({ta:.2f} s) and one frequency solve ({tf:.2f} s) of "
                      f"the same spiral; a CPU gradient by central differences takes "
                      f"{2 * s.n_params + 1} assemblies x (assembly + {len(freqs)} solves) = "
                      f"{per_grad:.1f} s on {threads} threads of {cpu_model()}"}
    return out


def plain_copy(mib: int, repeats: int = 5) -> dict:
    """A device-to-device copy of known size: FETCH_SIZE and WRITE_SIZE of
    its dispatches give the scale from the counters' units to bytes."""
    import torch

    n = mib * (1 << 20) // 4
    x = torch.rand(n, device="cuda:0", dtype=torch.float32)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(repeats):
        y.copy_(x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / repeats
    assert torch.equal(x, y)
    return {"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "repeats": repeats,
            "ms_per_copy": dt * 1e3, "gb_per_s": 8 * n / dt / 1e9}


def library_gemm_ceiling(n: int = 8192, reps: int = 10) -> dict:
    """The vendor library's bf16 GEMM (rocBLAS gemm_ex, bf16 inputs, FP32
    accumulation and output, n x n x n, random data) on the same GPU: what
    the bf16 matrix cores deliver in practice on this board, beside the
    published dense peak the roofline is priced against."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    blas = C.CDLL("librocblas.so")
    rng = np.random.default_rng(1)
    bufs = []

    def dev(nbytes, host=None):
        ptr = C.c_void_p()
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
        bufs.append(ptr)
        if host is not None:
            hip.hipMemcpy(ptr, host.ctypes.data_as(C.c_void_p), C.c_size_t(nbytes), 1)
        return ptr

    def bf16(shape):
        x = rng.standard_normal(shape, dtype=np.float32)
        return np.ascontiguousarray((x.view(np.uint32) >> 16).astype(np.uint16))

    a = dev(n * n * 2, bf16((n, n)))
    b = dev(n * n * 2, bf16((n, n)))
    c = dev(n * n * 4, np.zeros((n, n), dtype=np.float32))
    handle = C.c_void_p()
    blas.rocblas_create_handle(C.byref(handle))
    alpha, beta = C.c_float(1.0), C.c_float(0.0)
    bf16_t, f32_t, none = 168, 151, 111

    def gemm():
        st = blas.rocblas_gemm_ex(handle, none, none, n, n, n, C.byref(alpha), a, bf16_t, n, b, bf16_t,
                                  n, C.byref(beta), c, f32_t, n, c, f32_t, n, f32_t, 0, 0, 0)
        if st != 0:
            raise RuntimeError(f"rocblas_gemm_ex status {st}")

    try:
        for _ in range(3):
            gemm()
        hip.hipDeviceSynchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            gemm()
        hip.hipDeviceSynchronize()
        dt = (time.perf_counter() - t0) / reps
    finally:
        blas.rocblas_destroy_handle(handle)
        for ptr in bufs:
            hip.hipFree(ptr)
    tflops = 2.0 * n ** 3 / dt / 1e12
    return {"library": "rocBLAS gemm_ex bf16 -> fp32", "n": n, "tflops": tflops,
            "frac_of_peak": tflops / PEAKS_TFLOPS["bf16"]}


def model_pack(problem):
    from fasterhenry_amd import model

    return model.pack(*problem)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--bars", type=int, default=49)
    ap.add_argument("--freqs", type=int, default=41)
    ap.add_argument("--precision", default="fp32_refined", choices=["fp64", "fp32_refined"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the plain-FP64 sweep the refined result is compared with")
    ap.add_argument("--workload", default="dense", choices=["dense", "matrix-free", "batch", "adjoint"],
                    help="dense: the headline line (with the matrix-free, batched and adjoint sweeps "
                         "attached); matrix-free / batch / adjoint: only that workload, as its own line")
    ap.add_argument("--no-matrix-free", action="store_true",
                    help="leave the matrix-free sweep out of the headline line")
    ap.add_argument("--no-batch", action="store_true",
                    help="leave the batched sweep out of the headline line")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="leave the adjoint gradients out of the headline line")
    ap.add_argument("--no-fd", action="store_true",
                    help="adjoint workload: skip the central-difference check")
    ap.add_argument("--sweep-parts", type=int, default=0, choices=[0, 1, 2, 3, 4],
                    help="refined sweep: groups of frequencies factorized side by side on as many "
                         "streams (0: the autotuner's choice)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default schedule (outer block 256, two sweep parts)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (HBM counter scale for rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; run with --gpus 1",
                          "n_gpus": args.gpus}))
        return 2

    if args.copy_mib:
        print(json.dumps(plain_copy(args.copy_mib)))
        return 0

    from fasterhenry_amd import MeshSystem, fixtures, model, native

    if args.workload == "matrix-free":
        out = matrix_free(args.steps, args.warmup)
        print(json.dumps(out))
        return 0
    if args.workload == "batch":
        out = batch(args.steps, args.warmup)
        print(json.dumps(out))
        return 0
    if args.workload == "adjoint":
        out = adjoint(args.steps, args.warmup, fd_check=not args.no_fd, cpu=not args.no_cpu)
        print(json.dumps(out))
        return 0

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
    tuned, tune_s = None, 0.0
    if args.sweep_parts:
        system.set_schedule(system.schedule()["outer_block"], args.sweep_parts)
    elif not args.no_autotune:
        # the sweep schedule (outer block x sweep parts) from the project's
        # autotuner, timed on this system's own sweep before the warm-up
        t0 = time.perf_counter()
        tuned, best_ms = system.autotune(freqs)
        tune_s = time.perf_counter() - t0
        print(f"# autotuned schedule {tuned} ({best_ms:.1f} ms per sweep) in {tune_s:.1f} s", flush=True)

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
    # FP32 matrix cores, or the bf16 matrix cores for split-bf16 updates.
    # `achieved` counts the algorithmic flops only (8 real flops per complex
    # multiply-add of the trailing update, the FP32-equivalent work); the
    # three bf16 products that carry each split product are reported apart
    # as `matrix_core_tflops`, so frac <= 1/3 for split-bf16.
    arith = infos[-1]["factor_arith"]
    peak_kind = {0: "fp64", 1: "fp32", 2: "bf16"}[arith]
    peak = PEAKS_TFLOPS[peak_kind]
    achieved = upd_flops / (upd_ms * 1e-3) / 1e12 if upd_ms > 0 else 0.0
    unit_rate = upd_unit_flops / (upd_ms * 1e-3) / 1e12 if upd_ms > 0 else 0.0

    # The same update kernel alone: one more profiled sweep with the
    # frequencies in one group (no other stream's kernels beside its
    # launches), and the vendor library's bf16 GEMM on the same GPU.
    isolated = None
    if args.precision != "fp64":
        sched = system.schedule()
        system.set_graphs(False)
        system.set_schedule(sched["outer_block"], 1)
        system.sweep_raw(freqs)
        system.set_profiling(True)
        _, iso = system.sweep_raw(freqs)
        system.set_profiling(False)
        system.set_schedule(sched["outer_block"], sched["sweep_parts"])
        system.set_graphs(True)
        iso = iso.as_dict()
        iso_rate = iso["update_kernel_flops"] / (iso["update_kernel_ms"] * 1e-3) / 1e12
        iso_unit = iso["update_kernel_unit_flops"] / (iso["update_kernel_ms"] * 1e-3) / 1e12
        isolated = {"sweep": "one group of all frequencies on one stream", "achieved": iso_rate,
                    "frac": iso_rate / peak, "matrix_core_tflops": iso_unit,
                    "matrix_core_frac": iso_unit / peak,
                    "avg_launch_ms": iso["update_kernel_ms"] / max(iso["update_kernel_launches"], 1)}
        try:
            ceiling = library_gemm_ceiling()
            isolated["library_gemm"] = ceiling
            isolated["matrix_core_vs_library"] = iso_unit / ceiling["tflops"]
        except (OSError, RuntimeError) as e:
            isolated["library_gemm"] = {"error": str(e)}

    extra = {}
    m = counts.internal_meshes
    if args.precision != "fp64" and not args.no_compare:
        # the plain FP64 form on the same resident system, with the same
        # warm-up and step count (its workspace grows during the warm-up)
        system.set_precision("fp64")
        for w in range(args.warmup):
            system.sweep_raw(freqs)
        system.set_profiling(True)
        infos64 = []
        t1 = time.perf_counter()
        for k in range(args.steps):
            z64, info64 = system.sweep_raw(freqs)
            infos64.append(info64.as_dict())
            print(f"# fp64 step {k + 1}/{args.steps}: {info64.solve_s:.3f} s", flush=True)
        t64 = (time.perf_counter() - t1) / args.steps
        system.set_profiling(False)
        system.set_precision(args.precision)
        err = max(np.linalg.norm(z[k] - z64[k]) / np.linalg.norm(z64[k]) for k in range(len(freqs)))
        ms64 = sum(i["update_kernel_ms"] for i in infos64)
        n64 = sum(i["update_kernel_launches"] for i in infos64)
        fl64 = sum(i["update_kernel_flops"] for i in infos64)
        a64 = fl64 / (ms64 * 1e-3) / 1e12 if ms64 > 0 else 0.0
        extra = {"z_rel_err_vs_fp64": err, "fp64_solves_per_s": args.freqs / t64,
                 "speedup_vs_fp64": value / (args.freqs / t64),
                 "fp64_form": {"solves_per_s": args.freqs / t64, "ms_per_step": t64 * 1e3,
                               "steps": args.steps, "warmup": args.warmup,
                               "update_kernel": "right-looking trailing update (fp64 matrix cores)",
                               "update_achieved_tflops": a64, "update_frac": a64 / PEAKS_TFLOPS["fp64"],
                               "update_launches": n64, "update_avg_launch_ms": ms64 / max(n64, 1)}}

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
               "z_rel_err_gpu_vs_cpu": float(np.linalg.norm(zg - res["z"]) / np.linalg.norm(res["z"])),
               "nproc": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
               "cpu_model": cpu_model()}
        cpu["end_to_end"] = cpu_end_to_end(threads)

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
                   "refine_change_max": infos[-1]["refine_change_max"],
                   "schedule": system.schedule(), "autotuned": tuned is not None,
                   "autotune_s": tune_s},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak,
                     "flops": "algorithmic: 8 real flops per complex multiply-add of the update "
                              "(FP32-equivalent), (4/3) m^3 per frequency over all launches",
                     "traffic": UPDATE_TRAFFIC_BYTES_PER_FREQUENCY * args.freqs /
                     system.schedule()["sweep_parts"] if (args.bars, args.freqs, args.precision,
                                                           system.schedule()["outer_block"]) ==
                     (49, 41, "fp32_refined", 256) else None,
                     "kernel": "outer trailing update of the L D L^T factorization "
                               f"({peak_kind} matrix cores)",
                     "matrix_core_tflops": unit_rate,
                     "matrix_core_frac": unit_rate / peak,
                     "launches": upd_n, "avg_launch_ms": upd_ms / max(upd_n, 1),
                     "flops_per_launch": upd_flops / max(upd_n, 1),
                     "streams": system.schedule()["sweep_parts"] if args.precision != "fp64" else 1,
                     # with the parts' launches overlapping, a launch's own
                     # duration stretches; this is every update flop of the
                     # timed steps over their whole wall time
                     "device_achieved": upd_flops / elapsed / 1e12,
                     "device_frac": upd_flops / elapsed / 1e12 / peak,
                     "isolated": isolated},
        "cpu_baseline": cpu,
        **extra,
    }
    system.close()
    if not args.no_matrix_free:
        out["matrix_free"] = matrix_free(1, 1, dense_check=(packed, ports))
    if not args.no_batch:
        out["batch"] = batch(1, 1)
    if not args.no_adjoint:
        out["adjoint"] = adjoint(1, 1)
    print(json.dumps(out))
    system.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
