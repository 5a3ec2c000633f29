#!/usr/bin/env python3
"""Benchmark of the FBP hot path on one MI355X: batched parallel-beam filtered
backprojection at scale.

Workload (default): 2048 sinograms of 1800 angles x 2048 detector pixels from a
seeded 3-D phantom of 20 ellipsoids (analytic plane-section line integrals),
filtered with the Shepp-Logan-windowed ramp zero-padded to 4096 (fused
fft -> filter -> ifft kernel) and backprojected into a 2048^3 float32 volume.
One step = ramp filter of all 3,686,400 rows + backprojection of all 2048
slices; inputs are resident in HBM before the timed region.  value = voxel
updates per second (GUPS) over the whole step.

Prints one JSON line (rank 0).  Also measures, inside the timed region, the
average duration of the dominant kernel (the backprojector) with HIP events
on the stream it runs on, and expresses it against the LDS read bandwidth the
device sustains (measured here by a microbenchmark); the CPU baseline is the
oracle (C, OpenMP) on one of the 2048 slices.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ufo-filters_amd"))

MI355X_HBM_TBPS = 8.0
MI355X_CUS = 256
MI355X_CLOCK_GHZ = 2.4


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measure_lds_peak(ctx, torch):
    sink = torch.empty(4096 * 256, device="cuda")
    ctx.bench_lds_read(4096, 64, sink)  # warm-up
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    nbytes = ctx.bench_lds_read(4096, 4096, sink)
    e1.record()
    torch.cuda.synchronize()
    return nbytes / (e0.elapsed_time(e1) * 1e-3)


def measure_hbm_copy(ctx, torch):
    n = 1 << 28  # 1 GiB per buffer
    a = torch.ones(n, device="cuda")
    b = torch.empty(n, device="cuda")
    ctx.bench_copy(a, b)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        ctx.bench_copy(a, b)
    e1.record()
    torch.cuda.synchronize()
    bw = 5 * 2 * 4 * n / (e0.elapsed_time(e1) * 1e-3)
    del a, b
    return bw


def cpu_baseline(sinos, n_fft, axis, threads):
    """Oracle (C, float32 reference semantics, OpenMP) on a few slices: ramp chain + BP."""
    import numpy as np
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    n, p, w = sinos.shape
    coef = oracle.filter_coefficients(2 * n_fft, "ramp-fromreal")
    t0 = time.perf_counter()
    for s in sinos:
        filt = oracle.ramp_chain(s, n_fft, coef, w).astype(np.float32)
        oracle.bp_texture_ref32(filt, axis_pos=axis)
    dt = time.perf_counter() - t0
    return (n * w * w * p) / dt / 1e9, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--slices", type=int, default=2048)
    ap.add_argument("--width", type=int, default=2048)
    ap.add_argument("--angles", type=int, default=1800)
    ap.add_argument("--storage", choices=["f32", "f16", "bf16"], default="f32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: every box in this setup has a single MI355X; "
              "slices are independent, so N GPUs would each run this same workload on 2048/N slices.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import ufomi

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    ctx = ufomi.Context(0)
    ctx.set_stream(torch.cuda.current_stream(0))

    S, W, P = args.slices, args.width, args.angles
    N = 1
    while N < 2 * W:
        N *= 2
    axis = W / 2.0
    step_angle = math.pi / P
    dtype = {"f32": ufomi.F32, "f16": ufomi.F16, "bf16": ufomi.BF16}[args.storage]
    tdtype = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}[args.storage]

    # ---- synthetic input, resident in HBM
    log("generating %d sinograms %dx%d from 20 seeded ellipsoids" % (S, P, W))
    ellipsoids = ufomi.phantom_ellipsoids(1001, 20, W / 2.0)
    sino = torch.empty((S, P, W), device=dev)
    chunk = 256
    for s0 in range(0, S, chunk):
        n = min(chunk, S - s0)
        ells = np.stack([ufomi.slice_ellipsoids(ellipsoids, (s0 + k) + 0.5 - S / 2.0) for k in range(n)])
        ctx.sinogram_ellipses(ells, W, P, axis, step_angle, sino[s0:s0 + n])
    coef = torch.from_numpy(ufomi.filter_coefficients(2 * N, "shepp-logan")).to(dev)
    filt = torch.empty((S, P, W), dtype=tdtype, device=dev)
    vol = torch.empty((S, W, W), device=dev)
    torch.cuda.synchronize()
    log("inputs ready (%.1f GB sinograms, %.1f GB volume)" % (sino.numel() * 4 / 1e9, vol.numel() * 4 / 1e9))

    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]

    def step(events=None):
        if events:
            events[0].record()
        ctx.ramp_filter(sino, W, S * P, N, coef, W, filt, dtype)
        if events:
            events[1].record()
        ctx.bp_parallel(filt, W, P, vol, n_slices=S, dtype=dtype, axis_pos=axis)
        if events:
            events[2].record()

    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log("warmup step %d done" % i)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(ev[i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    filt_ms = [ev[i][0].elapsed_time(ev[i][1]) for i in range(args.steps)]
    bp_ms = [ev[i][1].elapsed_time(ev[i][2]) for i in range(args.steps)]
    log("timed %d steps: %.3f s wall; filter %.2f ms, backprojection %.1f ms per step"
        % (args.steps, wall, float(np.mean(filt_ms)), float(np.mean(bp_ms))))

    updates = float(S) * W * W * P
    ms_per_step = wall / args.steps * 1e3
    gups = updates / (wall / args.steps) / 1e9

    # ---- roofline of the dominant kernel (backprojector): 2 taps x sample bytes per update
    lds_peak = measure_lds_peak(ctx, torch)
    hbm_copy = measure_hbm_copy(ctx, torch)
    sample_bytes = 4 if args.storage == "f32" else 2
    bp_s = float(np.mean(bp_ms)) * 1e-3
    achieved = updates * 2 * 4 / bp_s  # LDS bytes: the interpolation reads (v, dv) float32 pairs
    filt_s = float(np.mean(filt_ms)) * 1e-3
    filt_bytes = float(S) * P * W * (4 + sample_bytes)
    log("backprojector %.1f GUPS, LDS gather %.1f TB/s of %.1f TB/s measured peak; "
        "filter %.2f TB/s HBM (algorithmic) of %.2f TB/s copy"
        % (updates / bp_s / 1e9, achieved / 1e12, lds_peak / 1e12, filt_bytes / filt_s / 1e12, hbm_copy / 1e12))

    result = {
        "metric": "GUPS (voxel*proj/s) for FBP",
        "value": round(gups, 2),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.storage == "f32" else "f32 accumulate, %s storage" % args.storage,
        "data": "synthetic: seeded 20-ellipsoid phantom, analytic parallel-beam sinograms",
        "config": {
            "workload": "batched parallel-beam FBP: %d sinograms x %d angles x %d px -> %dx%dx%d volume, "
                        "Shepp-Logan-windowed ramp zero-padded to %d, %s filtered-sinogram storage"
                        % (S, P, W, S, W, W, N, args.storage),
            "slices": S, "angles": P, "detector": W, "fft_size": N, "storage": args.storage,
        },
        "roofline": {
            "kernel": "bp_texture_kernel (4 slices per workgroup)",
            "bound": "lds",
            "achieved": round(achieved / 1e9, 1),
            "peak": round(lds_peak / 1e9, 1),
            "unit": "GB/s",
            "frac": round(achieved / lds_peak, 4),
            "traffic": None,
            "algorithmic_bytes_per_update": 8,
            "kernel_ms": round(bp_s * 1e3, 2),
            "kernel_gups": round(updates / bp_s / 1e9, 1),
            "peak_source": "measured: ufomi_bench_lds_read (ds_read_b128, all CUs)",
            "hbm_copy_GBps": round(hbm_copy / 1e9, 1),
            "filter_kernel_ms": round(filt_s * 1e3, 2),
            "filter_hbm_GBps": round(filt_bytes / filt_s / 1e9, 1),
            "filter_hbm_frac_of_8TBps": round(filt_bytes / filt_s / (MI355X_HBM_TBPS * 1e12), 4),
        },
    }

    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        n_cpu = min(8, S)
        some = sino[S // 2:S // 2 + n_cpu].cpu().numpy()
        del sino, filt, vol
        log("cpu baseline: oracle on %d slices (%d threads)" % (n_cpu, threads))
        cg, dt = cpu_baseline(some, N, axis, threads)
        result["cpu_baseline"] = {
            "value": round(cg, 3), "unit": "GUPS", "cores": threads, "kind": "port",
            "sample": "%d of %d slices (%dx%d px, %d angles each): FP64 fft-filter-ifft + float32-semantics "
                      "backprojection of the C oracle, %.1f s" % (n_cpu, S, W, W, P, dt),
        }
    print(json.dumps(result), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
