"""Benchmark of the gridding hot path on one MI355X.

Workload (the headline configuration): 3D radial kooshball, 51,200 spokes on
Fibonacci-sphere directions x 256 samples (13,107,200 samples per coil),
32 coils, complex64 seeded Gaussian data, DCF |k|^2, Kaiser-Bessel width 4,
oversampling 2 (128^3 image -> 256^3 grid), 512-entry LUT, FP32
accumulation.  Inputs are resident in HBM before the timed region.

A step is one gridding pass over all 32 coils (fr_grid_execute).  `value`
is gridded coil-samples per second = C * M / t_step.  The JSON line also
carries the roofline of the gridding pass (algorithmic bytes
C*M*8 + M*(4*3+4) + C*N^3*8 per pass over its average duration measured
with HIP events on the launch stream), the full reconstruction latency
(grid + cropped FFT + fused shift/crop/SOS), the plan build time, the other
configurations' legs, the parity of the timed passes' grid against the
oracle (FP32 restatement of the reference CPU loop and FP64), and the CPU
baseline: that FP32 restatement timed on this host over the whole workload
with min(coils, host threads) threads.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--spokes", type=int, default=51200)
    ap.add_argument("--readout", type=int, default=256)
    ap.add_argument("--coils", type=int, default=32)
    ap.add_argument("--image", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-recon", action="store_true")
    ap.add_argument("--no-ldcf", action="store_true", help="skip the learned-vs-iterative DCF leg")
    ap.add_argument("--no-frames", action="store_true", help="skip the batched 2D spiral frames leg and the 2D radial leg")
    ap.add_argument("--no-fp16", action="store_true", help="skip the FP16-sample 384^3 leg")
    ap.add_argument("--accumulate", type=int, default=0)
    ap.add_argument("--exact", action="store_true", help="exact-order plan for the timed steps")
    ap.add_argument("--unit", type=int, nargs=3, default=None, metavar=("X", "Y", "Z"),
                    help="base work-unit box (default: the plan's choice)")
    ap.add_argument("--split", type=int, default=0, help="slice/split threshold in taps (0: automatic)")
    ap.add_argument("--ipw", type=int, default=0, help="accumulate 6: light items per wave (0: automatic)")
    ap.add_argument("--xcd-run", type=int, default=0, help="accumulate 6: items per XCD run (0: automatic)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (the HBM-counter calibration kernel) and exit")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "grid_pos_kooshball3d_c32_traffic.txt"),
                    help="counter summary giving the gridding pass's HBM bytes (FETCH_SIZE + WRITE_SIZE)")
    return ap.parse_args()


def copy_calibration():
    """A plain device copy of known size: what FETCH_SIZE / WRITE_SIZE read
    for it calibrates those counters (bytes counted per byte moved)."""
    import torch
    n = 1 << 28                      # 2^28 floats = 1 GiB each way
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_each_way": 4 * n, "launches": 5}), flush=True)


GRID_KERNELS = {1: "k_grid_items (LDS atomic add)", 2: "k_grid_items (LDS read-add-write)",
                4: "k_grid_gather (cell-sorted batches, VGPR cells)", 5: "k_grid_mfma (matrix cores)",
                6: "k_grid_pos (matrix cores, weights from sample positions)"}


# sources whose code decides the gridding pass's memory traffic; a counter
# summary applies only to the build it was measured on
TRAFFIC_SOURCES = ("fastrecon_amd/csrc/fr_grid_pos.hip", "fastrecon_amd/csrc/fr_plan_pos.hip",
                   "fastrecon_amd/csrc/fr_grid.hip", "fastrecon_amd/csrc/fr_grid_mfma.hip",
                   "fastrecon_amd/csrc/fr_plan.hip")


def sources_hash():
    import hashlib
    h = hashlib.sha256()
    for rel in TRAFFIC_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def read_traffic(path, workload, kernel):
    """HBM bytes per gridding pass from a committed counter summary, used
    only when it was measured on this exact workload, kernel and build (the
    hash of the gridding sources recorded in the summary)."""
    try:
        kv = {}
        with open(path) as f:
            for line in f:
                if ":" in line and not line.startswith("#"):
                    k, v = line.split(":", 1)
                    kv[k.strip()] = v.strip()
        if kv.get("workload") != workload or kv.get("kernel") != kernel or kv.get("sources_sha16") != sources_hash():
            return None, None
        return float(kv["traffic_bytes_per_pass"]), os.path.relpath(path, ROOT)
    except (OSError, KeyError, ValueError):
        return None, None


def time_stream(torch, ctx, fn, steps):
    stream = torch.cuda.ExternalStream(ctx.stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    ctx.sync()
    return e0.elapsed_time(e1) / steps


def device_data(torch, ctx, samples, M, C, seed):
    """Seeded complex Gaussian coil data generated on the device, sample-major."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    d = torch.randn(M * C * 2, generator=gen, device="cuda", dtype=torch.float32)
    torch.cuda.synchronize()
    samples.set_data_device(d.data_ptr(), 1)
    ctx.sync()
    del d


def bench_frames(ctx, torch, hip, syn, steps):
    """64 uneven 2D spiral frames (8-48 interleaves x 4096 samples), 12 coils,
    192^2 image (384^2 grid): one batched plan, every frame gridded in one
    launch, the whole series (grid + batched FFT + per-frame shift/crop/SOS)
    replayed as one HIP graph; the autotuner's pick is timed too."""
    C, image = 12, 192
    N = syn.grid_size(image, 2.0)
    frames, counts = syn.spiral_frames(64, 4096)
    tns = [hip.normalize_traj(k, N)[0] for k, _ in frames]
    sizes = np.array([t.shape[0] for t in tns], np.int64)
    M = int(sizes.sum())
    s = hip.Samples(ctx, 2, M, C, frame_samples=sizes)
    s.set_traj(np.concatenate(tns))
    s.set_dcf(np.concatenate([d for _, d in frames]))
    device_data(torch, ctx, s, M, C, 4)
    lut, _ = hip.kb_lut(4.0, 2.0, 512)
    F = len(tns)
    res = {"workload": f"64 uneven 2D spiral frames ({int(counts.min())}-{int(counts.max())} interleaves x 4096), "
                       f"{C} coils, {image}^2 image, og 2 -> {N}^2 grid", "samples_total": M}
    t0 = time.perf_counter()
    best, best_ms, def_ms, timed = hip.autotune_search(ctx, s, N, 4.0, lut, reps=3)
    res["autotune"] = {"search_s": time.perf_counter() - t0, "timed_configs": timed, "search_default_ms": def_ms,
                       "search_best_ms": best_ms, "best": {k: v for k, v in best.items() if v}}
    for label, tune in (("default", None), ("tuned", best)):
        t0 = time.perf_counter()
        p = hip.Plan(ctx, s, N, 4.0, lut, tune)
        plan_ms = (time.perf_counter() - t0) * 1e3
        g = hip.Grids(ctx, 2, N, C, frames=F)
        fft = hip.FFT(ctx, 2, N, C * F, +1)
        img = hip.DeviceBuffer(ctx, 4 * image * image * F)
        graph = hip.Graph(p, s, g, fft, (image, image, 1), True, None, img)
        grid_ms = time_stream(torch, ctx, lambda: hip.grid_execute(p, s, g), steps)
        series_ms = time_stream(torch, ctx, graph.launch, steps)
        res[label] = {"plan_ms": plan_ms, "grid_ms": grid_ms, "series_ms_graph": series_ms,
                      "gridded_coil_samples_per_s": C * M / (grid_ms * 1e-3)}
        for h in (graph, img, fft, g, p):
            h.close()
    s.close()
    return res


def bench_fp16(ctx, torch, hip, syn, steps):
    """3D kooshball 100,000 spokes x 384, 16 coils, samples stored as FP16
    pairs (FP32 accumulation), 256^3 image, og 1.5 -> 384^3 grid: gridding,
    and the recon with the cropped N = 384 transforms, shift and coil sum."""
    C, image, og = 16, 256, 1.5
    N = syn.grid_size(image, og)
    traj, dcf = syn.kooshball3d(100000, 384)
    M = traj.shape[0]
    tn, _ = hip.normalize_traj(traj, N)
    del traj
    s = hip.Samples(ctx, 3, M, C, hip.FR_SAMPLES_C16)
    s.set_traj(tn)
    s.set_dcf(dcf)
    device_data(torch, ctx, s, M, C, 3)
    lut, _ = hip.kb_lut(4.0, og, 512)
    t0 = time.perf_counter()
    best, best_ms, def_ms, timed = hip.autotune_search(ctx, s, N, 4.0, lut, reps=2)
    tune_s = time.perf_counter() - t0
    g = hip.Grids(ctx, 3, N, C)
    f = hip.FFT(ctx, 3, N, C, +1)
    img = hip.DeviceBuffer(ctx, 4 * image ** 3)
    rc = (image, image, image)
    b = C * M * 4 + M * 16 + C * N ** 3 * 8
    res = {"workload": f"kooshball3d 100000 spokes x 384, {C} coils FP16, {image}^3 image, og 1.5 -> {N}^3 grid",
           "samples_per_coil": M, "algorithmic_bytes": b,
           "autotune": {"search_s": tune_s, "timed_configs": timed, "search_default_ms": def_ms,
                        "search_best_ms": best_ms, "best": {k: v for k, v in best.items() if v}}}
    for label, tune in (("default", None), ("tuned", best)):
        t0 = time.perf_counter()
        p = hip.Plan(ctx, s, N, 4.0, lut, tune)
        plan_ms = (time.perf_counter() - t0) * 1e3
        hip.grid_execute(p, s, g)
        hip.fft_sos(f, g, rc, True, None, img)
        ms = time_stream(torch, ctx, lambda: hip.grid_execute(p, s, g), steps)

        def recon():
            hip.grid_execute(p, s, g)
            hip.fft_sos(f, g, rc, True, None, img)
        recon_ms = time_stream(torch, ctx, recon, steps)
        res[label] = {"plan_ms": plan_ms, "grid_ms": ms, "recon_ms": recon_ms,
                      "gridded_coil_samples_per_s": C * M / (ms * 1e-3),
                      "achieved_gbs": b / (ms * 1e-3) / 1e9, "frac_of_8tbs": b / (ms * 1e-3) / 8e12}
        p.close()
    for h in (img, f, g, s):
        h.close()
    return res


def bench_radial2d(ctx, torch, hip, syn, steps):
    """2D golden-angle radial, 402 spokes x 512 samples, 8 coils, 256^2 image,
    og 2 -> 512^2 grid (the reference program's CPU-sized case): gridding and
    the recon with the cropped transforms, shift, crop and coil sum."""
    C, image = 8, 256
    N = syn.grid_size(image, 2.0)
    traj, dcf = syn.radial2d(402, 512)
    M = traj.shape[0]
    tn, _ = hip.normalize_traj(traj, N)
    s = hip.Samples(ctx, 2, M, C)
    s.set_traj(tn)
    s.set_dcf(dcf)
    device_data(torch, ctx, s, M, C, 4)
    lut, _ = hip.kb_lut(4.0, 2.0, 512)
    t0 = time.perf_counter()
    p = hip.Plan(ctx, s, N, 4.0, lut)
    plan_ms = (time.perf_counter() - t0) * 1e3
    g = hip.Grids(ctx, 2, N, C)
    f = hip.FFT(ctx, 2, N, C, +1)
    img = hip.DeviceBuffer(ctx, 4 * image * image)
    rc = (image, image, 1)
    hip.grid_execute(p, s, g)
    hip.fft_sos(f, g, rc, True, None, img)
    grid_ms = time_stream(torch, ctx, lambda: hip.grid_execute(p, s, g), steps)

    def recon():
        hip.grid_execute(p, s, g)
        hip.fft_sos(f, g, rc, True, None, img)

    recon_ms = time_stream(torch, ctx, recon, steps)
    b = C * M * 8 + M * 12 + C * N * N * 8
    res = {"workload": f"radial2d 402 spokes x 512, {C} coils, {image}^2 image, og 2 -> {N}^2 grid",
           "samples_per_coil": M, "plan_ms": plan_ms, "grid_ms": grid_ms, "recon_ms": recon_ms,
           "gridded_coil_samples_per_s": C * M / (grid_ms * 1e-3), "algorithmic_bytes": b,
           "achieved_gbs": b / (grid_ms * 1e-3) / 1e9}
    for h in (img, f, g, p, s):
        h.close()
    return res


def cpu_info():
    """Host threads available to this process, all online threads, CPU model."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return avail, os.cpu_count() or avail, model


def cpu_baseline_and_parity(torch, hip, traj_n, dcf, N, coils, lut, grid_dev, seed):
    """The oracle's FP32 restatement of GridLut::griddingChannel
    (GridLut.cpp:80-147) on the whole workload, OpenMP over coils with
    min(coils, host threads) threads as Program/Main.cpp:111 sets them,
    timed; then the device grid's relative L2 error against it and against
    the FP64 oracle (coil by coil), on the same coil data."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import concurrent.futures as cf
    avail, online, model = cpu_info()
    threads = max(1, min(coils, avail))
    M = traj_n.shape[0]
    # the device's coil data (same seed), to the host, coil-major
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    d = torch.randn(M * coils * 2, generator=gen, device="cuda", dtype=torch.float32)
    data = np.ascontiguousarray(d.view(M, coils, 2).cpu().numpy().view(np.complex64).reshape(M, coils).T)
    del d
    torch.cuda.empty_cache()
    t0 = time.perf_counter()
    ref = oracle.grid_f32(traj_n, dcf, data, N, 4.0, lut, nthreads=threads)
    dt = time.perf_counter() - t0
    got = grid_dev.download_all()
    num = den = num64 = num_cpu = 0.0
    for c in range(coils):
        num += float(np.sum(np.abs(got[c] - ref[c]).astype(np.float64) ** 2))
        den += float(np.sum(np.abs(ref[c]).astype(np.float64) ** 2))

    def f64(c):
        return c, oracle.grid_f64(traj_n, dcf, data[c:c + 1], N, 4.0, lut, nthreads=1)[0]
    den64 = 0.0
    with cf.ThreadPoolExecutor(threads) as ex:
        for c, r64 in ex.map(f64, range(coils)):
            num64 += float(np.sum(np.abs(got[c].astype(np.complex128) - r64) ** 2))
            num_cpu += float(np.sum(np.abs(ref[c].astype(np.complex128) - r64) ** 2))
            den64 += float(np.sum(np.abs(r64) ** 2))
    base = {"value": coils * M / dt, "unit": "coil-samples/s", "cores": threads, "kind": "port",
            "host_threads_available": avail, "host_threads_online": online, "cpu_model": model,
            "sample": f"the whole workload ({M} samples x {coils} coils on the {N}^3 grid), {dt:.2f} s, "
                      f"oracle FP32 restatement of GridLut.cpp:80-147, OpenMP over coils with min(coils, "
                      f"host threads) threads (Main.cpp:111)"}
    parity = {"rel_l2_vs_cpu_fp32": float(np.sqrt(num / den)), "rel_l2_vs_fp64": float(np.sqrt(num64 / den64)),
              "cpu_fp32_vs_fp64": float(np.sqrt(num_cpu / den64)), "bar": "rel-L2 <= 1e-5 vs the CPU FP32 path",
              "data": "the timed passes' coil data, every sample and coil"}
    return base, parity


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "design-only)"}))
        sys.exit(2)
    import torch
    from fastrecon_amd import hip
    from fastrecon_amd import synthetic as syn

    og, kw, L = 2.0, 4.0, 512
    N = syn.grid_size(a.image, og)
    C = a.coils
    traj, dcf = syn.kooshball3d(a.spokes, a.readout)
    M = traj.shape[0]
    lut, _ = hip.kb_lut(kw, og, L)
    traj_n, _ = hip.normalize_traj(traj, N)

    ctx = hip.Context(0)
    s = hip.Samples(ctx, 3, M, C)
    s.set_traj(traj_n)
    s.set_dcf(dcf)
    # seeded complex Gaussian data generated on the device, sample-major
    gen = torch.Generator(device="cuda")
    gen.manual_seed(2)
    data = torch.randn(M * C * 2, generator=gen, device="cuda", dtype=torch.float32)
    torch.cuda.synchronize()
    s.set_data_device(data.data_ptr(), hip.FR_LAYOUT_SAMPLE_MAJOR)
    ctx.sync()
    del data
    torch.cuda.empty_cache()

    t0 = time.perf_counter()
    tune = {"accumulate": a.accumulate, "exact_order": 1 if a.exact else 0, "split_taps": a.split,
            "items_per_wave": a.ipw, "xcd_run": a.xcd_run}
    if a.unit:
        tune.update(tile_x=a.unit[0], tile_y=a.unit[1], tile_z=a.unit[2])
    plan = hip.Plan(ctx, s, N, kw, lut, tune)
    plan_ms = (time.perf_counter() - t0) * 1e3
    info = plan.info()
    grids = hip.Grids(ctx, 3, N, C)

    stream = torch.cuda.ExternalStream(ctx.stream)
    for _ in range(a.warmup):
        hip.grid_execute(plan, s, grids)
    ctx.sync()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
    t_wall = time.perf_counter()
    for e0, e1 in ev:
        e0.record(stream)
        hip.grid_execute(plan, s, grids)
        e1.record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t_wall) * 1e3 / a.steps
    kms = [e0.elapsed_time(e1) for e0, e1 in ev]
    ms = float(np.mean(kms))

    bytes_alg = C * M * 8 + M * (4 * 3 + 4) + C * N ** 3 * 8
    achieved = bytes_alg / (ms * 1e-3) / 1e9
    workload = (f"kooshball3d: 3D radial {a.spokes} Fibonacci spokes x {a.readout} samples, "
                f"{C} coils, {a.image}^3 image, og 2 -> {N}^3 grid, KB width 4, L 512, complex64, DCF |k|^2")
    kernel = (GRID_KERNELS.get(info["accumulate"], "unknown")
              + f", unit {info['tile_x']}x{info['tile_y']}x{info['tile_z']}")
    traffic, traffic_src = read_traffic(a.traffic_file, workload, kernel) if not (
        a.exact or a.unit or a.split or a.accumulate or a.ipw or a.xcd_run) else (None, None)
    out = {
        "metric": "gridded samples/s, recon ms (3D radial, 32 coils); learned-DCF vs iterative",
        "value": C * M / (ms * 1e-3),
        "unit": "coil-samples/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "wall_ms_per_step": wall_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload,
                   "samples_per_coil": M, "coils": C, "grid": N,
                   "tile": [info["tile_x"], info["tile_y"], info["tile_z"]], "slabs": info["slabs"],
                   "accumulate": ("automatic: " if a.accumulate == 0 else "") + kernel,
                   "order": "exact (CPU summation order)" if a.exact else "sliced (fixed-order slice combine)",
                   "items": info["items"], "heavy_units": info["heavy_units"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "algorithmic_bytes": bytes_alg,
                     "kernel": f"gridding pass: {kernel} + k_combine (fr_grid_execute; k_pos_dcf only on the "
                               "first pass, the DCF being unchanged after it), HIP events on the context stream",
                     "sources_sha16": sources_hash()},
        "plan_ms": plan_ms, "plan_taps": info["taps"], "plan_bytes": info["plan_bytes"],
    }

    if not a.no_recon:
        fft = hip.FFT(ctx, 3, N, C, +1)
        img = hip.DeviceBuffer(ctx, 4 * a.image ** 3)
        rc = (a.image, a.image, a.image)
        hip.grid_execute(plan, s, grids); hip.fft_sos(fft, grids, rc, True, None, img)
        ctx.sync()
        r = []
        for _ in range(5):
            e0, e1, e2 = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
            e0.record(stream)
            hip.grid_execute(plan, s, grids)
            e1.record(stream)
            hip.fft_sos(fft, grids, rc, True, None, img)
            e2.record(stream)
            ctx.sync()
            r.append((e0.elapsed_time(e1), e1.elapsed_time(e2), e0.elapsed_time(e2)))
        r = np.median(np.array(r), axis=0)
        out["recon_ms"] = float(r[2])
        out["recon_stages_ms"] = {"grid": float(r[0]), "fft_shift_crop_sos": float(r[1])}
        # the same image stage as a full in-place batched FFT then the SOS epilogue
        r = []
        for _ in range(3):
            hip.grid_execute(plan, s, grids)
            e1, e2, e3 = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
            e1.record(stream)
            fft(grids)
            e2.record(stream)
            hip.sos(ctx, grids, rc, True, None, img)
            e3.record(stream)
            ctx.sync()
            r.append((e1.elapsed_time(e2), e2.elapsed_time(e3)))
        r = np.median(np.array(r), axis=0)
        out["recon_stages_ms"]["full_fft_then_sos"] = {"fft": float(r[0]), "shift_crop_sos": float(r[1])}
        fft.close()
        img.close()

    if not a.no_frames:
        grids.close()
        torch.cuda.empty_cache()
        out["frames_spiral2d"] = bench_frames(ctx, torch, hip, syn, a.steps)
        grids = hip.Grids(ctx, 3, N, C)
    if not a.no_fp16:
        out["fp16_kooshball3d"] = bench_fp16(ctx, torch, hip, syn, max(5, a.steps // 2))
    if not a.no_frames:
        out["radial2d"] = bench_radial2d(ctx, torch, hip, syn, a.steps)

    if not a.no_ldcf:
        # learned vs iterative DCF on this workload's geometry (the network's
        # checkpoint was trained on the device by fastrecon_amd.ldcf_train)
        from fastrecon_amd import ldcf
        grids.close()
        plan.close()
        s.close()
        torch.cuda.empty_cache()
        mlp, meta = ldcf.load(ctx)

        def fill(smp):
            gen2 = torch.Generator(device="cuda")
            gen2.manual_seed(2)
            d = torch.randn(M * C * 2, generator=gen2, device="cuda", dtype=torch.float32)
            torch.cuda.synchronize()
            smp.set_data_device(d.data_ptr(), hip.FR_LAYOUT_SAMPLE_MAJOR)
            ctx.sync()

        r = ldcf.evaluate(ctx, mlp, traj, fill, N, (a.image, a.image, a.image), ramp=dcf, reps=2, coils=C)
        out["ldcf"] = dict(r, iterations=10, checkpoint_rows=meta.get("rows"), checkpoint_loss=meta["losses"][-1][1])
        out["ldcf"]["phantom_single_coil"] = ldcf.evaluate_phantom(ctx, mlp, traj, N, (a.image,) * 3, a.image,
                                                                   ramp=dcf)
        mlp.close()
        # the same comparison on the 2D radial geometry (402 x 512, 256^2
        # image, 512^2 grid) with 32 coils and the 2D network
        mlp2, meta2 = ldcf.load(ctx, ldcf.CHECKPOINT_2D)
        t2, r2 = syn.radial2d(402, 512)
        out["ldcf_2d"] = dict(ldcf.evaluate(ctx, mlp2, t2, syn.coil_data(32, t2.shape[0], 5), ldcf.N_GRID_2D,
                                            (ldcf.IMAGE_2D, ldcf.IMAGE_2D, 1), ramp=r2, reps=3),
                              workload="radial2d 402 spokes x 512, 32 coils, 256^2 image, og 2 -> 512^2 grid",
                              checkpoint_loss=meta2["losses"][-1][1])
        mlp2.close()

    if not a.no_cpu:
        # the timed passes' grid, recomputed on the same data (the legs above
        # reused the device memory)
        s2 = hip.Samples(ctx, 3, M, C)
        s2.set_traj(traj_n)
        s2.set_dcf(dcf)
        device_data(torch, ctx, s2, M, C, 2)
        plan2 = hip.Plan(ctx, s2, N, kw, lut, tune)
        g2 = hip.Grids(ctx, 3, N, C)
        hip.grid_execute(plan2, s2, g2)
        ctx.sync()
        plan2.close()
        s2.close()
        out["cpu_baseline"], out["parity"] = cpu_baseline_and_parity(torch, hip, traj_n, dcf, N, C, lut, g2, 2)
        g2.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
