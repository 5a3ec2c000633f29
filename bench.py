#!/usr/bin/env python3
"""Benchmark of the FBP hot path on one MI355X.

Default workload (the headline metric): batched parallel-beam FBP at scale.
2048 sinograms of 1800 angles x 2048 detector pixels from a seeded 3-D phantom
of 20 ellipsoids (analytic plane-section line integrals) are filtered with the
Shepp-Logan-windowed ramp zero-padded to 4096 (fused fft -> filter -> ifft
kernel) and backprojected into a 2048^3 float32 volume.  One step = ramp
filter of all 3,686,400 rows + backprojection of all 2048 slices; inputs are
resident in HBM before the timed region.  value = voxel updates per second
(GUPS) over the whole step.

Other workloads (--workload):
  fbp512  one 512^2 slice, 512 angles, 1% noise, Ram-Lak zero-padded to 1024
  cone    FDK: cosine weight + ramp (N 2048) + cone-beam backprojection of 720
          projections of 1024^2 into 1024^3 (SOD 2500 px, SDD 5000 px)
  lamino  ramp (N 2048) + laminographic backprojection, tilt 30 deg, 1500
          projections of 1024^2 into 1024^2 x 256
  learned 512^2 textured phantoms: train the residual U-Net on (120-view FBP,
          1800-view FBP) pairs, then slices/s and RMSE to the phantom of 120-view
          FBP + U-Net against plain 1800-view FBP (--views 60 for 60 views)
  transpose  transpose-projections of 1800 projections of 2048^2 (the fbp2048
          data as projections) into 2048 sinograms: GB/s against HBM
  batch   256 single-slice requests of uneven ROI (log-uniform in [256,2048]^2)
          and angle count ({360,720,1200,1800,3600}) as one HIP graph: one
          segmented ramp launch per FFT size + one persistent backprojection
          launch, tile / angle chunk picked by the autotuner

Prints one JSON line.  The dominant kernel's average duration is measured
inside the timed region with HIP events on the stream it runs on and set
against the LDS read bandwidth the device sustains (microbenchmark); the HBM
traffic comes from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of
the same workload (profiles/traffic.json) when there is one.  The CPU baseline
is the oracle (C, OpenMP) on a bounded sample of the same workload.
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measure_lds_peak(ctx, torch):
    """Median of 7 runs of the conflict-free ds_read_b128 microbenchmark, each long enough (16384
    iterations on 8192 workgroups) that clock ramps and neighbours on the node average out."""
    sink = torch.empty(8192 * 256, device="cuda")
    ctx.bench_lds_read(8192, 256, sink)  # warm-up
    rates = []
    for _ in range(7):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        nbytes = ctx.bench_lds_read(8192, 16384, sink)
        e1.record()
        torch.cuda.synchronize()
        rates.append(nbytes / (e0.elapsed_time(e1) * 1e-3))
    rates.sort()
    log("LDS read peak runs (TB/s): %s" % " ".join("%.1f" % (v / 1e12) for v in rates))
    return rates[len(rates) // 2]


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


def traffic_entry(key):
    """Counter figures of one kernel and workload from profiles/traffic.json: HBM bytes per launch
    (rocprof FETCH_SIZE and WRITE_SIZE, scaled) and VALU instructions per voxel update (SQ_INSTS_VALU)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f).get(key, {})


def traffic_for(key):
    entry = traffic_entry(key)
    return entry["fetch_bytes"] + entry["write_bytes"] if "fetch_bytes" in entry else None


# VALU issue capacity: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD every 4 cycles of 2.4 GHz,
# i.e. 16 lane-operations per SIMD and cycle
VALU_LANE_OPS = 256 * 4 * 16 * 2.4e9


def want_accuracy(args):
    """The FP64 error probes run with the CPU baseline, and alone for the secondary workloads."""
    return not args.no_cpu_baseline or getattr(args, "accuracy_only", False)


def want_cpu(args):
    return not args.no_cpu_baseline and not getattr(args, "accuracy_only", False)


# The reference's own OpenCL kernels cannot serve as the CPU baseline here: the host's OpenCL
# platform (AMD APP, clinfo) lists the GPU only and no CPU ICD is installed (probe summary in
# profiles/host_opencl_probe.md).  The baseline is the float32 C restatement of those kernels.
CPU_NOTE = ("CPU OpenCL baseline unavailable: the host's only OpenCL platform (AMD APP) lists the GPU alone, "
            "no CPU ICD (probed with clinfo); reported instead: the float32 C restatement of the reference's "
            "kernels, OpenMP on the host's cores")


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def oracle_module(threads):
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    return oracle


def run_timed(torch, steps, warmup, step_fn, n_events):
    """warmup untimed steps, then `steps` timed steps; returns wall seconds and per-step event lists."""
    for i in range(warmup):
        step_fn(None)
        torch.cuda.synchronize()
        log("warmup step %d done" % i)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(n_events)] for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        step_fn(ev[i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, ev


def mean_ms(ev, a, b):
    return sum(e[a].elapsed_time(e[b]) for e in ev) / len(ev)


# ----------------------------------------------------------------- workloads

STORAGE_NAMES = {"f32": "float", "f16": "__half", "bf16": "__hip_bfloat16"}


def fbp_parallel(args, ctx, torch, np, ufomi, S, W, P, filt_name, noise, storages):
    """Batched parallel-beam FBP (the default workload at S=2048, W=2048, P=1800).

    Every storage in `storages` (filtered-sinogram type: f32, f16, bf16) runs the same
    filter + backprojection steps on the same resident sinograms; the first one is the
    reported value, the others are reported beside it."""
    dev = torch.device("cuda:0")
    N = 1
    while N < 2 * W:
        N *= 2
    axis = W / 2.0
    step_angle = math.pi / P
    probe = S // 2  # the slice compared with the FP64 oracle

    log("generating %d sinograms %dx%d" % (S, P, W))
    sino = torch.empty((S, P, W), device=dev)
    if S == 1:
        ell = ufomi.phantom_ellipses(1000, 10, W / 2.0)
        ctx.sinogram_ellipses(ell, W, P, axis, step_angle, sino)
    else:
        ellipsoids = ufomi.phantom_ellipsoids(1001, 20, W / 2.0)
        chunk = 256
        for s0 in range(0, S, chunk):
            n = min(chunk, S - s0)
            ells = np.stack([ufomi.slice_ellipsoids(ellipsoids, (s0 + k) + 0.5 - S / 2.0) for k in range(n)])
            ctx.sinogram_ellipses(ells, W, P, axis, step_angle, sino[s0:s0 + n])
    if noise:
        ctx.add_noise(sino, noise, 1000)
    coef = torch.from_numpy(ufomi.filter_coefficients(2 * N, filt_name)).to(dev)
    vol = torch.empty((S, W, W), device=dev)
    torch.cuda.synchronize()
    log("inputs ready (%.2f GB sinograms, %.2f GB volume)" % (sino.numel() * 4 / 1e9, vol.numel() * 4 / 1e9))

    runs = {}
    for storage in storages:
        dtype = {"f32": ufomi.F32, "f16": ufomi.F16, "bf16": ufomi.BF16}[storage]
        tdtype = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}[storage]
        filt = torch.empty((S, P, W), dtype=tdtype, device=dev)

        def step(e):
            if e: e[0].record()
            ctx.ramp_filter(sino, W, S * P, N, coef, W, filt, dtype)
            if e: e[1].record()
            ctx.bp_parallel(filt, W, P, vol, n_slices=S, dtype=dtype, axis_pos=axis)
            if e: e[2].record()

        log("storage %s" % storage)
        wall, ev = run_timed(torch, args.steps, args.warmup, step, 3)
        filt_ms, bp_ms = mean_ms(ev, 0, 1), mean_ms(ev, 1, 2)
        log("timed %d steps: %.3f s wall; filter %.2f ms, backprojection %.2f ms per step"
            % (args.steps, wall, filt_ms, bp_ms))
        runs[storage] = dict(wall=wall, filt_ms=filt_ms, bp_ms=bp_ms, slice=vol[probe].cpu().numpy())
        del filt
        torch.cuda.empty_cache()

    updates = float(S) * W * W * P
    first = storages[0]

    def kname(st):
        # the instance ufomi_bp_parallel picks: slices per tile, angle chunk, half staging
        sk = 8 if S >= 8 else (4 if S >= 4 else (2 if S >= 2 else 1))
        return "bp_texture_kernel<%d,%d,%s,%s>" % (sk, 8 if sk >= 4 else 16, STORAGE_NAMES[st],
                                                   "true" if st == "f16" and sk >= 4 else "false")
    kernel = kname(first)
    sample_bytes = 4 if first == "f32" else 2
    filt_ms = runs[first]["filt_ms"]
    extra = {
        "filter_kernel_ms": round(filt_ms, 3),
        "filter_hbm_GBps": round(S * P * W * (4 + sample_bytes) / (filt_ms * 1e-3) / 1e9, 1),
        "filter_hbm_frac_of_8TBps": round(S * P * W * (4 + sample_bytes) / (filt_ms * 1e-3) / 8e12, 4),
    }
    variants = {st: {"value": round(updates / (r["wall"] / args.steps) / 1e9, 2),
                     "ms_per_step": round(r["wall"] / args.steps * 1e3, 3),
                     "bp_kernel_ms": round(r["bp_ms"], 3), "filter_kernel_ms": round(r["filt_ms"], 3),
                     "bp_kernel": kname(st)}
                for st, r in runs.items()}
    cpu = accuracy = None
    if want_accuracy(args):
        threads = min(16, os.cpu_count() or 1)
        n_cpu = min(8 if S > 1 else 1, S)
        some = sino[probe:probe + n_cpu].cpu().numpy()
        del sino, vol
        torch.cuda.empty_cache()
        oracle = oracle_module(threads)
        # error of every storage against the FP64 oracle (FP64 fft-filter-ifft + FP64 backprojection)
        t0 = time.perf_counter()
        ref = oracle.bp_texture_f64(oracle.ramp_chain(some[0], N, oracle.filter_coefficients(2 * N, filt_name), W),
                                    axis_pos=axis)
        accuracy = {"vs": "FP64 C oracle (FP64 %s fft-filter-ifft + FP64 backprojection) of slice %d"
                          % (filt_name, probe)}
        for st, r in runs.items():
            variants[st]["rel_l2_vs_fp64"] = float(oracle.rel_l2(r["slice"], ref))
            accuracy[st] = variants[st]["rel_l2_vs_fp64"]
        log("FP64 oracle slice %.1f s; rel-L2 %s" % (time.perf_counter() - t0, accuracy))
    if want_cpu(args):
        log("cpu baseline: oracle on %d slices (%d threads)" % (n_cpu, threads))
        coef_o = oracle.filter_coefficients(2 * N, filt_name)
        reps = 1 if S > 1 else 10
        t0 = time.perf_counter()
        for _ in range(reps):
            for s in some:
                f32 = oracle.ramp_chain_f32(s, N, coef_o, W)
                oracle.bp_texture_ref32(f32, axis_pos=axis)
        dt = time.perf_counter() - t0
        cpu = {"value": round(reps * n_cpu * W * W * P / dt / 1e9, 3), "unit": "GUPS", "cores": threads,
               "host_cores": host_cores(), "kind": "port", "note": CPU_NOTE,
               "sample": "%d of %d slices (%dx%d px, %d angles each)%s: float32 fft-filter-ifft + float32 "
                         "backprojection (backproject.cl semantics) of the C restatement, %.1f s"
                         % (n_cpu, S, W, W, P, " x %d repeats" % reps if reps > 1 else "", dt)}
    workload = ("parallel-beam FBP: %d sinogram%s x %d angles x %d px -> %dx%dx%d volume, %s ramp zero-padded "
                "to %d, %s filtered-sinogram storage" % (S, "s" if S > 1 else "", P, W, S, W, W, filt_name, N, first))
    if len(storages) > 1:
        extra["storage_variants"] = variants
    r = runs[first]
    # LDS bytes read per update: (v, d) float pairs, or half pairs with f16 half staging
    bpu = 4 if "true" in kernel else 8
    # compulsory HBM bytes of one backprojection launch: the stored filtered sinograms once, the volume once
    hbm_bytes = float(S) * P * W * (4 if first == "f32" else 2) + float(S) * W * W * 4
    return dict(updates=updates, wall=r["wall"], bp_ms=r["bp_ms"], kernel=kernel, bytes_per_update=bpu,
                hbm_bytes=hbm_bytes, extra=extra, cpu=cpu, accuracy=accuracy, workload=workload,
                dtype="f32" if first == "f32" else "f32 accumulate, %s storage" % first,
                config={"slices": S, "angles": P, "detector": W, "fft_size": N, "storage": first, "filter": filt_name},
                traffic_key="%s %dx%dx%d %s" % (kernel, S, P, W, first),
                metric="GUPS (voxel*proj/s) for FBP")


def general(args, ctx, torch, np, ufomi, kind):
    """Cone-beam FDK (kind='cone') or laminography (kind='lamino') through general-backproject."""
    dev = torch.device("cuda:0")
    W = H = 1024
    N = 2048
    if kind == "cone":
        P, nz = 720, 1024
        sod = 2500.0
        geom = dict(source_position_y=[-sod], detector_position_y=[sod], center_position_x=[W / 2.0],
                    center_position_z=[H / 2.0])
    else:
        P, nz = 1500, 256
        geom = dict(axis_angle_x=[math.radians(30.0)], center_position_x=[W / 2.0], center_position_z=[H / 2.0])
    params = ufomi.gbp_params(P, region=(-nz / 2.0, nz / 2.0, 1.0))
    g = ufomi.gbp_geometry(**geom)
    rows, _ = ufomi.gbp_projection_rows(g, params)
    ell = ufomi.phantom_ellipsoids(1003, 20, W / 2.0 * (0.9 if kind == "cone" else 0.5))
    if kind == "lamino":
        ell[:, 2] *= 0.25  # a flat sample, as laminography expects
        ell[:, 5] *= 0.25
    projs = torch.empty((P, H, W), device=dev)
    ctx.project_ellipsoids(ell, W, H, rows, projs)
    weighted = torch.empty_like(projs)
    filt = torch.empty_like(projs)
    coef = torch.from_numpy(ufomi.filter_coefficients(2 * N, "ramp-fromreal")).to(dev)
    vol = torch.empty((nz, W, W), device=dev)
    torch.cuda.synchronize()
    log("%s: %d projections %dx%d -> volume %dx%dx%d" % (kind, P, H, W, nz, W, W))

    def step(e):
        if e: e[0].record()
        src = projs
        if kind == "cone":
            ctx.cone_weight(projs, weighted, W, H, P, W / 2.0, H / 2.0, sod, sod, 0.0)
            src = weighted
        ctx.ramp_filter(src, W, P * H, N, coef, W, filt)
        if e: e[1].record()
        ctx.gbp(filt, W, H, P, g, params, vol)
        if e: e[2].record()

    wall, ev = run_timed(torch, args.steps, args.warmup, step, 3)
    pre_ms, bp_ms = mean_ms(ev, 0, 1), mean_ms(ev, 1, 2)
    log("timed %d steps: %.3f s wall; weight+filter %.2f ms, backprojection %.2f ms per step"
        % (args.steps, wall, pre_ms, bp_ms))
    updates = float(nz) * W * W * P
    # the kernel instance gbp_launch picks: z-separable maps (no z term in U, D, C) take the 2-D path
    zsep = bool(np.all(rows[:, 0, 2] == 0) and (kind != "cone" or (np.all(rows[:, 2, 2] == 0)
                                                               and np.all(rows[:, 3, 2] == 0))))
    mode = (1 if zsep else 2) if kind == "cone" else (3 if zsep else 0)
    kernel = "gbp_lds_kernel<%d,2,float>" % mode
    cpu = accuracy = None
    if want_accuracy(args):
        threads = min(16, os.cpu_count() or 1)
        oracle = oracle_module(threads)
        # probe: the central 64 x 64 x 4 voxels and 256 detector rows of the filter
        probe_z = 4
        c = W // 2
        got_sub = vol[nz // 2:nz // 2 + probe_z, c - 32:c + 32, c - 32:c + 32].cpu().numpy()
        rows_idx = np.linspace(0, P * H - 1, 256).astype(np.int64)
        src_rows = (weighted if kind == "cone" else projs).reshape(P * H, W)[rows_idx].cpu().numpy()
        got_rows = filt.reshape(P * H, W)[rows_idx].cpu().numpy()
        f_np = filt.cpu().numpy()
        del projs, weighted, filt, vol
        torch.cuda.empty_cache()
        t0 = time.perf_counter()
        ref_rows = oracle.ramp_chain(src_rows, N, oracle.filter_coefficients(2 * N, "ramp-fromreal"), W)
        ref_sub = oracle.gbp(f_np, overall_angle=2 * math.pi, region=(0.0, float(probe_z), 1.0),
                             x_region=(-32.0, 32.0, 1.0), y_region=(-32.0, 32.0, 1.0), **geom)
        accuracy = {"vs": "FP64 C oracle: FP64 fft-filter-ifft of 256 detector rows; FP64 general backprojection "
                          "of the central 64x64x%d voxels from the device-filtered projections" % probe_z,
                    "filter_rel_l2": float(oracle.rel_l2(got_rows, ref_rows)),
                    "backprojection_rel_l2": float(oracle.rel_l2(got_sub, ref_sub))}
        log("FP64 oracle probe %.1f s: %s" % (time.perf_counter() - t0, accuracy))
    if want_cpu(args):
        sub = 64 if kind == "cone" else 16
        x_reg = (-128.0, 128.0, 1.0)
        t0 = time.perf_counter()
        oracle.gbp(f_np, overall_angle=2 * math.pi, region=(0.0, float(sub), 1.0), x_region=x_reg,
                   y_region=x_reg, mode="ref32", **geom)
        dt = time.perf_counter() - t0
        n_up = sub * 256 * 256 * P
        cpu = {"value": round(n_up / dt / 1e9, 4), "unit": "GUPS", "cores": threads, "host_cores": host_cores(),
               "kind": "port", "note": CPU_NOTE,
               "sample": "central 256x256x%d voxels, all %d projections, float32-semantics general "
                         "backprojection of the C oracle, %.1f s" % (sub, P, dt)}
    desc = ("cone-beam FDK: %d projections of %dx%d, SOD %.0f px, SDD %.0f px, cosine weight + Ram-Lak (N %d) -> "
            "%dx%dx%d volume" % (P, H, W, sod, 2 * sod, N, nz, W, W) if kind == "cone" else
            "laminography: %d projections of %dx%d, tilt 30 deg, Ram-Lak (N %d) -> %dx%dx%d volume"
            % (P, H, W, N, nz, W, W))
    return dict(updates=updates, wall=wall, bp_ms=bp_ms, kernel=kernel, bytes_per_update=4 * 4,
                hbm_bytes=float(P) * H * W * 4 + float(nz) * W * W * 4,
                extra={"weight_filter_ms": round(pre_ms, 3)}, cpu=cpu, accuracy=accuracy, workload=desc, dtype="f32",
                config={"projections": P, "detector": [H, W], "volume": [nz, W, W], "fft_size": N},
                traffic_key="%s %s" % (kernel, kind),
                metric="GUPS (voxel*proj/s) for %s" % ("FDK cone-beam" if kind == "cone" else "laminography"))


def transpose(args, ctx, torch, np, ufomi):
    """transpose-projections of P projections [P][H][W] into H sinograms [H][P][W]; one step = one launch."""
    dev = torch.device("cuda:0")
    P, H, W = 1800, 2048, 2048
    g = torch.Generator(device=dev).manual_seed(17)
    projs = torch.randn((P, H, W), device=dev, generator=g)
    sinos = torch.empty((H, P, W), device=dev)
    torch.cuda.synchronize()
    log("transpose: %d projections of %dx%d (%.1f GiB each way)" % (P, H, W, projs.numel() * 4 / 2 ** 30))

    def step(e):
        if e: e[0].record()
        ctx.transpose_projections(projs, W, H, P, sinos)
        if e: e[1].record()

    wall, ev = run_timed(torch, args.steps, args.warmup, step, 2)
    k_ms = mean_ms(ev, 0, 1)
    moved = 2.0 * P * H * W * 4
    # exactness on a probe: sinogram rows equal projection rows
    probe = [(0, 0), (H - 1, P - 1), (H // 3, P // 2)]
    exact = all(torch.equal(sinos[h, p], projs[p, h]) for h, p in probe)
    cpu = None
    if not args.no_cpu_baseline:
        oracle = oracle_module(1)
        sub = 48
        host = projs[:, :sub].cpu().numpy()
        t0 = time.perf_counter()
        oracle.transpose_projections(host, P)
        dt = time.perf_counter() - t0
        cpu = {"value": round(2.0 * host.nbytes / dt / 1e9, 2), "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": "rows 0..%d of every projection (%d x %d x %d float32) through the numpy oracle, %.2f s"
                         % (sub - 1, P, sub, W, dt)}
    hbm_copy = measure_hbm_copy(ctx, torch)
    key = "transpose_rows_kernel transpose"
    traffic = traffic_for(key)
    result = {
        "metric": "GB/s transpose-projections (projections -> sinograms, bytes read + written)",
        "value": round(moved / (k_ms * 1e-3) / 1e9, 1), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic: seeded normal projections",
        "config": {"workload": "transpose-projections: %d projections of %dx%d -> %d sinograms of %dx%d"
                               % (P, H, W, H, P, W)},
        "roofline": {"kernel": "transpose_rows_kernel<float4>", "bound": "hbm",
                     "achieved": round(moved / (k_ms * 1e-3) / 1e9, 1), "peak": MI355X_HBM_TBPS * 1e3,
                     "unit": "GB/s", "frac": round(moved / (k_ms * 1e-3) / (MI355X_HBM_TBPS * 1e12), 4),
                     "traffic": traffic, "algorithmic_bytes": moved, "kernel_ms": round(k_ms, 3),
                     "hbm_copy_GBps": round(hbm_copy / 1e9, 1)},
        "exact": exact,
    }
    if cpu is not None:
        result["cpu_baseline"] = cpu
    return result


def batch(args, ctx, torch, np, ufomi):
    """256 uneven single-slice FBP requests as one HIP graph: segmented ramp filters + one persistent
    backprojection launch, tile / angle chunk chosen by the autotuner."""
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(2048)
    n = 256
    lo, hi = math.log(256), math.log(2048)
    reqs, shapes = [], []
    log("generating %d requests" % n)
    for i in range(n):
        rw = int(round(math.exp(rng.uniform(lo, hi))))
        rh = int(round(math.exp(rng.uniform(lo, hi))))
        P = int(rng.choice([360, 720, 1200, 1800, 3600]))
        W = max(rw, rh)
        sino = torch.empty((P, W), device=dev)
        ctx.sinogram_ellipses(ufomi.phantom_ellipses(3000 + i, 10, W / 2.0), W, P, W / 2.0, math.pi / P, sino)
        out = torch.empty((rh, rw), device=dev)
        reqs.append(dict(sinogram=sino, slice=out, roi_x=(W - rw) // 2, roi_y=(W - rh) // 2, roi_width=rw,
                         roi_height=rh))
        shapes.append((W, P, rw, rh))
    b = ufomi.FbpBatch(ctx, reqs, filter="ramp-fromreal")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tile, chunk, tune_ms = b.autotune(reps=2, force=True)
    tune_s = time.perf_counter() - t0
    info = b.info()
    updates = info["voxel_updates"]
    log("autotuner: tile %d, angle chunk %d (%.2f ms backprojection), %.2f s; %d items on %d workgroups, "
        "%d filter launches, %.3g voxel updates" % (tile, chunk, tune_ms, tune_s, info["items"], info["workgroups"],
                                                    info["filter_launches"], updates))

    def step(e):
        if e: e[0].record()
        b.run()
        if e: e[1].record()

    wall, ev = run_timed(torch, args.steps, args.warmup, step, 2)
    graph_ms = mean_ms(ev, 0, 1)
    # the persistent kernel and the filters launched directly, for the kernel's own duration
    ev2 = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    for e in ev2:
        e[0].record()
        b.launch(1)
        e[1].record()
        b.launch(2)
        e[2].record()
    torch.cuda.synchronize()
    filt_ms, bp_ms = mean_ms(ev2, 0, 1), mean_ms(ev2, 1, 2)
    log("graph %.2f ms per replay; direct launches: filters %.2f ms, persistent backprojection %.2f ms"
        % (graph_ms, filt_ms, bp_ms))
    cpu = accuracy = None
    cost = [rw * rh * P for (W, P, rw, rh) in shapes]
    if want_accuracy(args):
        threads = min(16, os.cpu_count() or 1)
        oracle = oracle_module(threads)
        b.run()
        torch.cuda.synchronize()
        # FP64 oracle on the two cheapest requests; the CPU baseline on requests in index order up to ~30 G updates
        cheap = sorted(range(n), key=lambda i: cost[i])[:2]
        accuracy = {"vs": "FP64 C oracle (FP64 ramp-fromreal fft-filter-ifft + FP64 backprojection) of requests %s"
                          % cheap}
        worst = 0.0
        for i in cheap:
            W, P, rw, rh = shapes[i]
            N = 1
            while N < 2 * W:
                N *= 2
            r = reqs[i]
            f64 = oracle.ramp_chain(r["sinogram"].cpu().numpy(), N, oracle.filter_coefficients(2 * N), W)
            ref = oracle.bp_texture_f64(f64, roi_x=r["roi_x"], roi_y=r["roi_y"], roi_width=rw, roi_height=rh)
            worst = max(worst, float(oracle.rel_l2(r["slice"].cpu().numpy(), ref)))
        accuracy["max_rel_l2"] = worst
        log("accuracy %s" % accuracy)
    if want_cpu(args):
        sample, acc_up = [], 0.0
        for i in range(n):
            if acc_up >= 30e9:
                break
            sample.append(i)
            acc_up += cost[i]
        host = [(reqs[i]["sinogram"].cpu().numpy(), shapes[i], reqs[i]) for i in sample]
        t0 = time.perf_counter()
        for s_np, (W, P, rw, rh), r in host:
            N = 1
            while N < 2 * W:
                N *= 2
            f32 = oracle.ramp_chain_f32(s_np, N, oracle.filter_coefficients(2 * N), W)
            oracle.bp_texture_ref32(f32, roi_x=r["roi_x"], roi_y=r["roi_y"], roi_width=rw, roi_height=rh)
        dt = time.perf_counter() - t0
        cpu = {"value": round(acc_up / dt / 1e9, 3), "unit": "GUPS", "cores": threads, "host_cores": host_cores(),
               "kind": "port", "note": CPU_NOTE,
               "sample": "requests 0..%d of %d (%.3g voxel updates): float32 fft-filter-ifft + float32 "
                         "backprojection of the C restatement, %.1f s" % (len(sample) - 1, n, acc_up, dt)}
    b.close()
    kernel = "bp_batch_kernel<%d,%d,float>" % (tile, chunk)
    extra = {"graph_ms": round(graph_ms, 3), "filter_kernels_ms": round(filt_ms, 3),
             "autotune": {"tile": tile, "angle_chunk": chunk, "tuned_bp_ms": round(tune_ms, 3),
                          "seconds": round(tune_s, 2)},
             "items": info["items"], "workgroups": info["workgroups"], "filter_launches": info["filter_launches"]}
    hbm_bytes = float(sum(P * W * 4 + rw * rh * 4 for W, P, rw, rh in shapes))
    return dict(updates=updates, wall=wall, bp_ms=bp_ms, kernel=kernel, bytes_per_update=8, hbm_bytes=hbm_bytes,
                extra=extra, cpu=cpu,
                accuracy=accuracy, dtype="f32",
                workload="uneven batched FBP: 256 requests, ROI log-uniform in [256,2048]^2, angles from "
                         "{360,720,1200,1800,3600}, detector width max(ROI), Ram-Lak zero-padded to 2^k >= 2W, "
                         "one HIP graph (segmented ramp launches + one persistent backprojection launch)",
                config={"requests": n, "voxel_updates": updates, "seed": 2048},
                traffic_key="%s batch" % kernel,
                metric="GUPS (voxel*proj/s) for uneven batched FBP")


def unet_flops(size):
    """Multiply-add flops (x2) of one U-Net forward pass on a size^2 image."""
    f = 2 * size * size * 9 * 1 * 32  # first conv
    for lev in range(4):
        hw = (size >> lev) ** 2
        C = 32 << lev
        if lev > 0:
            f += 2 * hw * 9 * (C // 2) * C
        f += 2 * hw * 9 * C * C
    for lev in range(3):
        hw = (size >> lev) ** 2
        C = 32 << lev
        f += 2 * (hw // 4) * (2 * C) * 4 * C  # transposed conv
        f += 2 * hw * 9 * (2 * C) * C + 2 * hw * 9 * C * C
    return f + 2 * size * size * 32


def unet_hbm_bytes(size):
    """HBM bytes of one U-Net forward pass on a size^2 image with every activation written once and
    read once by each consumer (no halo re-reads): the network's compulsory traffic."""
    P = [(size >> lev) ** 2 for lev in range(4)]
    C = [32 << lev for lev in range(4)]
    b = 4 * P[0] + 2 * 32 * P[0] + 2 * P[0]  # first conv: float image in, e1 and bf16 image out
    for lev in range(4):
        if lev > 0:
            b += 2 * C[lev - 1] * P[lev] + 2 * C[lev] * P[lev]  # first conv of the level: pooled in, e1 out
        b += 2 * 2 * C[lev] * P[lev]  # second conv: e1 in, skip half (or bottom) out
        if lev < 3:
            b += 2 * C[lev] * P[lev] // 4  # its 2x2 max pool
    for lev in range(2, -1, -1):
        b += 2 * 2 * C[lev] * P[lev + 1] + 2 * C[lev] * P[lev]  # transposed conv: source in, up half out
        b += 2 * 2 * C[lev] * P[lev] + 2 * C[lev] * P[lev]  # decoder conv a: concatenation in, d1 out
        b += 2 * C[lev] * P[lev] + (2 * C[lev] * P[lev] if lev > 0 else 8 * P[0])  # conv b (level 0: + image, output)
    return b


def learned_fbp(args, ctx, torch, np, ufomi):
    """Learned sparse-view FBP of 512^2 textured phantoms: train the U-Net on pairs
    (sparse-view FBP, 1800-view FBP), then compare slices/s and RMSE to the phantom of
    (sparse FBP + U-Net) against plain 1800-view FBP."""
    import learned
    dev = torch.device("cuda:0")
    size, full, sparse = 512, 1800, args.views
    n_train, n_test, batch = args.train_pairs, 256, 64
    steps = args.train_steps if args.train_steps > 0 else max(1, args.epochs * (n_train // batch))
    pm = learned.PairMaker(ctx, size, sparse, full, chunk=256)
    t0 = time.perf_counter()
    X = torch.empty((n_train, size, size), device=dev)
    Y = torch.empty_like(X)
    pm.make(np.arange(n_train) + 100000, X, Y)
    Xt = torch.empty((n_test, size, size), device=dev)
    Yt, Pt = torch.empty_like(Xt), torch.empty_like(Xt)
    pm.make(np.arange(n_test) + 900000, Xt, Yt, Pt)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    log("generated %d training pairs + %d test phantoms (%d / %d views) in %.1f s"
        % (n_train, n_test, sparse, full, gen_s))

    net = ufomi.Unet(ctx, batch, size, size, training=True, seed=2024)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = learned.train(net, X, Y, steps=steps, batch=batch, lr=1e-3, log=log, log_every=max(1, steps // 10))
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    log("trained %d steps (batch %d) in %.1f s" % (steps, batch, train_s))
    params = net.get_params()
    net.close()
    del X, Y
    torch.cuda.empty_cache()

    inf = ufomi.Unet(ctx, n_test, size, size)
    inf.set_params(params)
    out = torch.empty_like(Xt)
    learned.infer(inf, Xt, out, n_test)
    torch.cuda.synchronize()
    mask = learned.disk_mask(size)
    acc = {"rmse_to_phantom": {"plain_%d_view_fbp" % full: learned.rmse(Yt, Pt, mask),
                               "plain_%d_view_fbp" % sparse: learned.rmse(Xt, Pt, mask),
                               "learned_%d_view" % sparse: learned.rmse(out, Pt, mask)},
           "rmse_to_%d_view_fbp" % full: {"plain_%d_view_fbp" % sparse: learned.rmse(Xt, Yt, mask),
                                          "learned_%d_view" % sparse: learned.rmse(out, Yt, mask)},
           "test_set": "%d unseen phantoms, RMSE inside the reconstruction disk" % n_test}
    log("accuracy %s" % acc)

    # throughput: 256 slices end to end, sinograms resident in HBM
    B = n_test
    ell_w = [learned.phantom_params(900000 + i, size) for i in range(B)]
    sino_full = torch.empty((B, full, size), device=dev)
    ctx.sinogram_phantom(np.stack([e for e, _, _ in ell_w]), np.stack([w for _, w, _ in ell_w]), size / 2.0, size,
                         full, size / 2.0, math.pi / full, sino_full)
    sino_sparse = sino_full[:, ::full // sparse].contiguous()
    filt_full = torch.empty_like(sino_full)
    filt_sparse = torch.empty_like(sino_sparse)
    rec = torch.empty((B, size, size), device=dev)
    rec2 = torch.empty_like(rec)
    N = learned.fft_size(size)
    coef = torch.from_numpy(ufomi.filter_coefficients(2 * N, "ramp-fromreal")).to(dev)

    def plain(e):
        if e: e[0].record()
        ctx.ramp_filter(sino_full, size, B * full, N, coef, size, filt_full)
        ctx.bp_parallel(filt_full, size, full, rec, n_slices=B, axis_pos=size / 2.0)
        if e: e[1].record()

    def learned_step(e):
        if e: e[0].record()
        ctx.ramp_filter(sino_sparse, size, B * sparse, N, coef, size, filt_sparse)
        ctx.bp_parallel(filt_sparse, size, sparse, rec, n_slices=B, axis_pos=size / 2.0)
        if e: e[1].record()
        inf.forward(rec, rec2, batch=B)
        if e: e[2].record()

    wall_p, ev_p = run_timed(torch, args.steps, args.warmup, plain, 2)
    wall_l, ev_l = run_timed(torch, args.steps, args.warmup, learned_step, 3)
    fbp_sparse_ms, unet_ms = mean_ms(ev_l, 0, 1), mean_ms(ev_l, 1, 2)
    plain_sps = B / (wall_p / args.steps)
    learned_sps = B / (wall_l / args.steps)
    log("plain %d-view FBP %.0f slices/s; learned %d-view %.0f slices/s (FBP %.2f ms + U-Net %.2f ms per %d)"
        % (full, plain_sps, sparse, learned_sps, fbp_sparse_ms, unet_ms, B))
    sink = torch.empty(4096 * 256, device=dev)
    ctx.bench_mfma(1024, 64, sink)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    mf = ctx.bench_mfma(4096, 2048, sink)
    e1.record()
    torch.cuda.synchronize()
    mfma_peak = mf / (e0.elapsed_time(e1) * 1e-3)
    flops = unet_flops(size) * B
    achieved = flops / (unet_ms * 1e-3)
    hbm_bytes = unet_hbm_bytes(size) * B
    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        oracle = oracle_module(threads)
        import unet_oracle
        named = inf.unpack(params)
        cnet = unet_oracle.UNet({k: v for k, v in named.items()}, bf16_storage=False)
        n_cpu = 2
        s_np = sino_sparse[:n_cpu].cpu().numpy()
        coef_o = oracle.filter_coefficients(2 * N, "ramp-fromreal")
        t0 = time.perf_counter()
        for i in range(n_cpu):
            f32 = oracle.ramp_chain(s_np[i], N, coef_o, size).astype(np.float32)
            img = oracle.bp_texture_ref32(f32, axis_pos=size / 2.0)
            cnet.forward(img[None].astype(np.float64))
        dt = time.perf_counter() - t0
        cpu = {"value": round(n_cpu / dt, 4), "unit": "slices/s", "cores": threads, "kind": "port",
               "sample": "%d slices: C-oracle %d-view FBP + numpy float64 U-Net forward (same weights), %.1f s"
                         % (n_cpu, sparse, dt)}
    inf.close()
    extra = {"kernel_ms": round(unet_ms, 3), "algorithmic_bytes_per_forward": hbm_bytes,
             "flops_per_forward": flops, "flop_per_byte": round(flops / hbm_bytes, 1),
             "mfma_achieved_TFLOPs": round(achieved / 1e12, 1), "mfma_frac_of_2.5PFLOPs": round(achieved / 2.5e15, 4),
             "mfma_peak_measured_TFLOPs": round(mfma_peak / 1e12, 1), "fbp_sparse_ms": round(fbp_sparse_ms, 3)}
    return dict(learned_sps=learned_sps, plain_sps=plain_sps, wall=wall_l, kernel_ms=unet_ms, achieved=achieved,
                hbm_achieved=hbm_bytes / (unet_ms * 1e-3),
                extra=extra, cpu=cpu, accuracy=acc,
                training={"pairs": n_train, "steps": steps, "batch": batch, "lr": 1e-3, "optimizer": "Adam",
                          "seconds": round(train_s, 1), "data_seconds": round(gen_s, 1),
                          "loss_first_last": [losses[0][1], losses[-1][1]]},
                workload="learned sparse-view FBP, 512^2 textured ellipse phantoms: %d-view Ram-Lak FBP + residual "
                         "U-Net (4 levels, 32 base channels, 1.93M params, BF16 MFMA, FP32 accumulation), batch %d, "
                         "against plain %d-view FBP" % (sparse, B, full))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["fbp2048", "fbp512", "cone", "lamino", "batch", "learned", "transpose"],
                    default="fbp2048")
    ap.add_argument("--slices", type=int, default=2048)
    ap.add_argument("--storage", choices=["f32", "f16", "bf16", "all"], default=None,
                    help="filtered-sinogram storage; 'all' (the fbp2048 default) reports f32 as the value "
                         "and f16 / bf16 beside it")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="fbp2048: leave out the secondary configs (512^2 slice, uneven batch, cone, laminography)")
    ap.add_argument("--bp-config", default="", help="S,A,H: slices per tile, angle chunk, f16 half staging "
                                                     "(ufomi_bp_configure; default automatic)")
    ap.add_argument("--views", type=int, default=120, help="learned: sparse view count (divides 1800)")
    ap.add_argument("--train-pairs", type=int, default=20000, help="learned: training pairs")
    ap.add_argument("--train-steps", type=int, default=0, help="learned: Adam steps (0: --epochs epochs)")
    ap.add_argument("--epochs", type=int, default=5, help="learned: epochs over the training pairs")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: every box in this setup has a single MI355X; "
              "slices are independent, so N GPUs would each run this same workload on 2048/N slices.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import ufomi

    torch.cuda.set_device(0)
    ctx = ufomi.Context(0)
    ctx.set_stream(torch.cuda.current_stream(0))
    if args.bp_config:
        ctx.bp_configure(*[int(v) for v in args.bp_config.split(",")])

    storage = args.storage or ("all" if args.workload == "fbp2048" else "f32")
    storages = ["f32", "f16", "bf16"] if storage == "all" else [storage]
    if args.workload == "fbp2048":
        r = fbp_parallel(args, ctx, torch, np, ufomi, args.slices, 2048, 1800, "shepp-logan", 0.0, storages)
    elif args.workload == "fbp512":
        r = fbp_parallel(args, ctx, torch, np, ufomi, 1, 512, 512, "ramp-fromreal", 0.01, storages)
    elif args.workload == "batch":
        r = batch(args, ctx, torch, np, ufomi)
    elif args.workload == "transpose":
        print(json.dumps(transpose(args, ctx, torch, np, ufomi)), flush=True)
        ctx.close()
        return
    elif args.workload == "learned":
        r = learned_fbp(args, ctx, torch, np, ufomi)
        result = {
            "metric": "slices/s learned %d-view vs plain 1800-view FBP" % args.views,
            "value": round(r["learned_sps"], 1), "unit": "slices/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(r["wall"] / args.steps * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "bf16 (U-Net, fp32 accumulation); f32 FBP",
            "data": "synthetic: seeded ellipse phantoms with band-limited texture, analytic sinograms",
            "config": {"workload": r["workload"]},
            "plain_1800_view_slices_per_s": round(r["plain_sps"], 1),
            "learned_vs_plain": round(r["learned_sps"] / r["plain_sps"], 3),
            # 18.8 TFLOP over 73.4 GB of compulsory traffic is 256 flop/B, below the 312 flop/B at which
            # 2.5 PFLOP/s and 8 TB/s meet: the forward pass is bound by HBM, and its MFMA rate is reported beside it
            "roofline": dict({"kernel": "U-Net forward (conv3_t8_kernel, gemm_kernel<...,UpFwd> + VALU layers)", "bound": "hbm",
                              "achieved": round(r["hbm_achieved"] / 1e9, 1), "peak": MI355X_HBM_TBPS * 1e3, "unit": "GB/s",
                              "frac": round(r["hbm_achieved"] / (MI355X_HBM_TBPS * 1e12), 4),
                              "traffic": traffic_for("U-Net forward learned512 batch256")}, **r["extra"]),
            "accuracy": r["accuracy"], "training": r["training"],
        }
        if r["cpu"] is not None:
            result["cpu_baseline"] = r["cpu"]
        print(json.dumps(result), flush=True)
        ctx.close()
        return
    else:
        r = general(args, ctx, torch, np, ufomi, args.workload)

    lds_peak = measure_lds_peak(ctx, torch)
    hbm_copy = measure_hbm_copy(ctx, torch)
    roof = roofline_of(r, lds_peak, hbm_copy)
    variants = r["extra"].pop("storage_variants", None)
    roof.update(r["extra"])
    result = {
        "metric": r["metric"],
        "value": round(r["updates"] / (r["wall"] / args.steps) / 1e9, 2),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["wall"] / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": r["dtype"],
        "data": "synthetic: seeded ellipse/ellipsoid phantoms, analytic projections",
        "config": dict({"workload": r["workload"]}, **r["config"]),
        "roofline": roof,
    }
    if variants:
        result["storage_variants"] = variants
    if r.get("accuracy"):
        result["accuracy"] = r["accuracy"]
    if r["cpu"] is not None:
        result["cpu_baseline"] = r["cpu"]
    if args.workload == "fbp2048" and not args.no_other_configs:
        result["other_configs"] = other_configs(args, ctx, torch, np, ufomi, lds_peak, hbm_copy)
    print(json.dumps(result), flush=True)
    ctx.close()


def roofline_of(r, lds_peak, hbm_copy):
    """Roofline of a workload's dominant backprojection kernel: the LDS gather rate against the
    measured ds_read_b128 peak (algorithmic bytes per update x updates / kernel time), the VALU
    issue rate the counters imply, and the HBM-compulsory and counted HBM rates."""
    bp_s = r["bp_ms"] * 1e-3
    gups = r["updates"] / bp_s
    achieved = r["updates"] * r["bytes_per_update"] / bp_s
    entry = traffic_entry(r["traffic_key"])
    traffic = traffic_for(r["traffic_key"])
    log("%s: %.1f GUPS kernel, gather %.1f TB/s vs %.1f TB/s LDS peak, HBM copy %.2f TB/s"
        % (r["kernel"], gups / 1e9, achieved / 1e12, lds_peak / 1e12, hbm_copy / 1e12))
    roof = {
        "kernel": r["kernel"],
        "bound": "lds",
        "achieved": round(achieved / 1e9, 1),
        "peak": round(lds_peak / 1e9, 1),
        "unit": "GB/s",
        "frac": round(achieved / lds_peak, 4),
        "traffic": traffic,
        "algorithmic_bytes_per_update": r["bytes_per_update"],
        "kernel_ms": round(r["bp_ms"], 3),
        "kernel_gups": round(gups / 1e9, 1),
        "peak_source": "measured: ufomi_bench_lds_read (ds_read_b128 on every CU), median of 7 runs",
        "hbm_copy_GBps": round(hbm_copy / 1e9, 1),
        # the HBM-compulsory fraction beside the on-chip one: inputs and output once per launch
        "hbm_compulsory_bytes": r["hbm_bytes"],
        "hbm_compulsory_GBps": round(r["hbm_bytes"] / bp_s / 1e9, 1),
        "hbm_compulsory_frac_of_8TBps": round(r["hbm_bytes"] / bp_s / 8e12, 5),
    }
    if traffic:
        roof["hbm_achieved_GBps"] = round(traffic / bp_s / 1e9, 1)
    valu = entry.get("valu_insts_per_update")
    if valu:
        # what binds these gathers: VALU issue.  Instructions per update from SQ_INSTS_VALU (profiles/)
        roof["valu_insts_per_update"] = valu
        roof["valu_ceiling_gups"] = round(VALU_LANE_OPS / valu / 1e9, 1)
        roof["valu_issue_frac"] = round(gups * valu / VALU_LANE_OPS, 4)
    return roof


def other_configs(args, ctx, torch, np, ufomi, lds_peak, hbm_copy):
    """Configs 0, 2 and 3 after the headline one: a few timed steps each, the kernel's rooflines and the
    FP64 error probe (no CPU baselines, which `--workload <name>` runs report)."""
    import copy
    sub = copy.copy(args)
    sub.steps, sub.warmup, sub.accuracy_only, sub.no_cpu_baseline = min(args.steps, 5), min(args.warmup, 2), True, True
    out = {}
    torch.cuda.empty_cache()
    for name in ("fbp512", "batch", "cone", "lamino"):
        log("secondary workload %s" % name)
        t0 = time.perf_counter()
        if name == "fbp512":
            r = fbp_parallel(sub, ctx, torch, np, ufomi, 1, 512, 512, "ramp-fromreal", 0.01, ["f32"])
        elif name == "batch":
            r = batch(sub, ctx, torch, np, ufomi)
        else:
            r = general(sub, ctx, torch, np, ufomi, name)
        roof = roofline_of(r, lds_peak, hbm_copy)
        out[name] = {
            "workload": r["workload"], "metric": r["metric"],
            "value": round(r["updates"] / (r["wall"] / sub.steps) / 1e9, 2), "unit": "GUPS",
            "steps": sub.steps, "ms_per_step": round(r["wall"] / sub.steps * 1e3, 3),
            "roofline": {k: roof[k] for k in ("kernel", "bound", "achieved", "peak", "unit", "frac", "traffic",
                                              "kernel_ms", "kernel_gups", "valu_insts_per_update",
                                              "valu_ceiling_gups", "valu_issue_frac", "hbm_compulsory_frac_of_8TBps")
                         if k in roof},
            "accuracy": r.get("accuracy"),
            "seconds": round(time.perf_counter() - t0, 1),
        }
        torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    main()
