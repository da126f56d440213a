"""Parity at the full workloads of the benchmark configurations (c1-c5).

Each test runs the HIP path through the C ABI on the whole workload and
compares it with the oracle on the host on the same inputs:

* c1: 2D golden-angle radial, 402 x 512, 8 coils, 512^2 grid;
* c2: 3D kooshball, 51,200 Fibonacci spokes x 256, 32 coils, 256^3 grid;
* c3: 3D kooshball, 100,000 spokes x 384, 16 coils stored as FP16, 384^3 grid
  (og 1.5), FP64 error per coil on the dequantised inputs;
* c4: 64 uneven spiral frames (8-48 interleaves x 4096), 12 coils, 384^2;
* c5: the shipped learned-DCF network on c2's geometry with 32 coils.

Bars (the reference CPU path is oracle/fr_oracle.c's restatement of
GridLut.cpp:80-147 and Main.cpp:116-181):
* gridding rel-L2 <= 1e-5 against the FP32 loop, and no further from the FP64
  oracle than twice the FP32 loop is (+ 1e-7);
* exact-order plans bit-identical to the FP32 loop;
* images rel-L2 <= 1e-5.  The image oracle is the oracle's FP32 grid, a
  complex128 FFT (scipy's: the oracle's naive FP64 transform, pinned against
  numpy in test_oracle_image.py, would take minutes at 256^3), the oracle's
  fftshift and its FP32 coil-order sum of squares.

Coil data: complex Gaussian, sigma 1 per component, numpy PCG64 seeded per
coil.  The tests print the FP64 error of every configuration."""
import concurrent.futures as cf
import os
import time

import numpy as np
import pytest
import scipy.fft as sfft

from fastrecon_amd import hip
from fastrecon_amd import synthetic as syn

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(1200)]
TOL = 1e-5
THREADS = 16          # the box's CPU share for one GPU


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-30))


def coils(C, M, seed):
    """(C, M) complex64 seeded Gaussian (sigma 1 per component), one PCG64
    stream per coil, generated in parallel."""
    out = np.empty((C, M), np.complex64)

    def fill(c):
        rng = np.random.default_rng([seed, c])
        v = out[c].view(np.float32)
        rng.standard_normal(out=v, dtype=np.float32)
    with cf.ThreadPoolExecutor(THREADS) as ex:
        list(ex.map(fill, range(C)))
    return out


def log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = hip.Context(0)
    yield c
    c.close()


def grid_gpu(ctx, tn, dcf, data, N, lut, tune=None, sample_type=hip.FR_SAMPLES_C64, image=None, s=None):
    """Grid on the device; optionally also the cropped-transform image."""
    M, dim = tn.shape
    C = data.shape[0]
    own = s is None
    if own:
        s = hip.Samples(ctx, dim, M, C, sample_type)
        s.set_traj(tn)
        s.set_dcf(dcf)
        s.set_coils(data)
    t0 = time.perf_counter()
    p = hip.Plan(ctx, s, N, 4.0, lut, tune)
    plan_s = time.perf_counter() - t0
    g = hip.Grids(ctx, dim, N, C)
    hip.grid_execute(p, s, g)
    grid = g.download_all()
    img = None
    if image is not None:
        f = hip.FFT(ctx, dim, N, C, +1)
        img = hip.fft_sos(f, g, image, True).to_host(np.float32)
        f.close()
    info = p.info()
    for h in (g, p) + ((s,) if own else ()):
        h.close()
    return grid, img, info, plan_s


def oracle_image(orc, dim, N, grid, rc):
    """grid (C, N^dim) complex64 -> unnormalised FFT (+1) in complex128, cast
    to complex64, fftshift, crop + FP32 sum of squares in coil order."""
    C = grid.shape[0]
    shape = (N,) * dim
    spec = np.empty_like(grid)
    for c in range(C):
        f = sfft.ifftn(grid[c].reshape(shape).astype(np.complex128), workers=THREADS) * float(N ** dim)
        spec[c] = f.astype(np.complex64).ravel()
    spec = orc.fftshift(dim, N, spec)
    return orc.sos(dim, N, spec, rc)


# ------------------------------------------------------------------ c1
def test_c1_radial2d_full(ctx, orc):
    N, rc = 512, (256, 256, 1)
    traj, dcf = syn.radial2d(402, 512)
    tn, _ = hip.normalize_traj(traj, N)
    data = coils(8, tn.shape[0], 1)
    lut, _ = hip.kb_lut(4.0, 2.0, 512)
    got, img, info, _ = grid_gpu(ctx, tn, dcf, data, N, lut, image=rc)
    ref = orc.grid_f32(tn, dcf, data, N, 4.0, lut, nthreads=8)
    ref64 = orc.grid_f64(tn, dcf, data, N, 4.0, lut, nthreads=8)
    e32, e64, c64 = rel_l2(got, ref), rel_l2(got, ref64), rel_l2(ref, ref64)
    ex, _, _, _ = grid_gpu(ctx, tn, dcf, data, N, lut, tune=dict(exact_order=1))
    ref_img = oracle_image(orc, 2, N, ref, rc)
    ei = rel_l2(img, ref_img)
    log(f"[c1] mode {info['accumulate']} grid rel-L2 vs FP32 {e32:.2e}, vs FP64 {e64:.2e} (FP32 loop vs FP64 "
        f"{c64:.2e}); exact plan bitwise {np.array_equal(ex.view(np.int32), ref.view(np.int32))}; image {ei:.2e}")
    assert e32 <= TOL and e64 <= 2 * c64 + 1e-7
    np.testing.assert_array_equal(ex.view(np.int32), ref.view(np.int32))
    assert ei <= TOL


# ------------------------------------------------------------------ c2
@pytest.fixture(scope="module")
def c2(orc):
    N = 256
    t0 = time.perf_counter()
    traj, dcf = syn.kooshball3d(51200, 256)
    tn, _ = hip.normalize_traj(traj, N)
    del traj
    data = coils(32, tn.shape[0], 2)
    lut, _ = hip.kb_lut(4.0, 2.0, 512)
    t1 = time.perf_counter()
    ref = orc.grid_f32(tn, dcf, data, N, 4.0, lut, nthreads=THREADS)
    t2 = time.perf_counter()
    log(f"[c2] inputs {t1 - t0:.1f} s, oracle FP32 grid {t2 - t1:.1f} s on {THREADS} threads")
    return dict(N=N, tn=tn, dcf=dcf, data=data, lut=lut, ref=ref)


def test_c2_kooshball_grid_full(ctx, orc, c2):
    N, tn, dcf, data, lut, ref = (c2[k] for k in ("N", "tn", "dcf", "data", "lut", "ref"))
    got, _, info, plan_s = grid_gpu(ctx, tn, dcf, data, N, lut)
    e32 = rel_l2(got, ref)
    # FP64 oracle coil by coil (8.6 GB at once otherwise)
    num = den = num_cpu = 0.0

    def f64(c):
        return orc.grid_f64(tn, dcf, data[c:c + 1], N, 4.0, lut, nthreads=1)[0]
    with cf.ThreadPoolExecutor(THREADS) as ex:
        for c, r64 in enumerate(ex.map(f64, range(32))):
            num += float(np.sum(np.abs(got[c].astype(np.complex128) - r64) ** 2))
            num_cpu += float(np.sum(np.abs(ref[c].astype(np.complex128) - r64) ** 2))
            den += float(np.sum(np.abs(r64) ** 2))
    e64, c64 = np.sqrt(num / den), np.sqrt(num_cpu / den)
    log(f"[c2] 51200x256, 32 coils, 256^3: mode {info['accumulate']} unit {info['tile_x']}x{info['tile_y']}x"
        f"{info['tile_z']} plan {plan_s * 1e3:.0f} ms; rel-L2 vs FP32 loop {e32:.2e}, vs FP64 {e64:.2e} "
        f"(FP32 loop vs FP64 {c64:.2e})")
    assert e32 <= TOL
    assert e64 <= 2 * c64 + 1e-7


def test_c2_kooshball_exact_order_bitwise(ctx, orc, c2):
    N, tn, dcf, data, lut, ref = (c2[k] for k in ("N", "tn", "dcf", "data", "lut", "ref"))
    got, _, info, _ = grid_gpu(ctx, tn, dcf, data, N, lut, tune=dict(exact_order=1))
    same = np.array_equal(got.view(np.int32), ref.view(np.int32))
    log(f"[c2] exact-order plan (mode {info['accumulate']}): bitwise equal to the FP32 loop: {same}")
    assert same


def test_c2_kooshball_image_full(ctx, orc, c2):
    N, tn, dcf, data, lut, ref = (c2[k] for k in ("N", "tn", "dcf", "data", "lut", "ref"))
    rc = (128, 128, 128)
    _, img, _, _ = grid_gpu(ctx, tn, dcf, data, N, lut, image=rc)
    ref_img = oracle_image(orc, 3, N, ref, rc)
    e = rel_l2(img, ref_img)
    log(f"[c2] image (grid, cropped transforms, shift, coil SOS) rel-L2 vs oracle {e:.2e}")
    assert e <= TOL


# ------------------------------------------------------------------ c3
def test_c3_fp16_kooshball_full(ctx, orc):
    N = 384
    traj, dcf = syn.kooshball3d(100000, 384)
    tn, _ = hip.normalize_traj(traj, N)
    del traj
    M, C = tn.shape[0], 16
    data = coils(C, M, 3)
    lut, _ = hip.kb_lut(4.0, 1.5, 512)
    s = hip.Samples(ctx, 3, M, C, hip.FR_SAMPLES_C16)
    s.set_traj(tn)
    s.set_dcf(dcf)
    s.set_coils(data)
    # the oracle grids the dequantised FP16 values
    deq = np.empty_like(data)
    deq.real = data.real.astype(np.float16).astype(np.float32)
    deq.imag = data.imag.astype(np.float16).astype(np.float32)
    del data
    p = hip.Plan(ctx, s, N, 4.0, lut)
    info = p.info()
    g = hip.Grids(ctx, 3, N, C)
    hip.grid_execute(p, s, g)

    def both(c):
        return (orc.grid_f32(tn, dcf, deq[c:c + 1], N, 4.0, lut, nthreads=1)[0],
                orc.grid_f64(tn, dcf, deq[c:c + 1], N, 4.0, lut, nthreads=1)[0])
    e32, e64, c64 = [], [], []
    num = den = 0.0
    with cf.ThreadPoolExecutor(THREADS) as ex:
        for c, (r32, r64) in enumerate(ex.map(both, range(C))):
            gc = g.download(c)
            e32.append(rel_l2(gc, r32))
            e64.append(rel_l2(gc.astype(np.complex128), r64))
            c64.append(rel_l2(r32.astype(np.complex128), r64))
            num += float(np.sum(np.abs(gc - r32) ** 2))
            den += float(np.sum(np.abs(r32) ** 2))
    for h in (g, p, s):
        h.close()
    log(f"[c3] 100000x384, 16 FP16 coils, 384^3: mode {info['accumulate']}; rel-L2 vs FP32 loop {np.sqrt(num / den):.2e}")
    log("[c3] per-coil rel-L2 vs FP64 oracle (dequantised inputs): " + " ".join(f"{x:.2e}" for x in e64))
    log("[c3] per-coil FP32 loop vs FP64:                          " + " ".join(f"{x:.2e}" for x in c64))
    assert np.sqrt(num / den) <= TOL
    for a, b in zip(e64, c64):
        assert a <= 2 * b + 1e-7


# ------------------------------------------------------------------ c4
def test_c4_spiral_frames_full(ctx, orc):
    C, image = 12, 192
    N = syn.grid_size(image, 2.0)
    frames, counts = syn.spiral_frames(64, 4096)
    tns = [hip.normalize_traj(k, N)[0] for k, _ in frames]
    dcfs = [d for _, d in frames]
    sizes = np.array([t.shape[0] for t in tns], np.int64)
    M = int(sizes.sum())
    F = len(tns)
    data = coils(C, M, 4)
    s = hip.Samples(ctx, 2, M, C, frame_samples=sizes)
    s.set_traj(np.concatenate(tns))
    s.set_dcf(np.concatenate(dcfs))
    s.set_coils(data)
    lut, _ = hip.kb_lut(4.0, 2.0, 512)
    p = hip.Plan(ctx, s, N, 4.0, lut)
    info = p.info()
    g = hip.Grids(ctx, 2, N, C, frames=F)
    hip.grid_execute(p, s, g)
    got = g.download_all().reshape(C, F, N * N)
    fft = hip.FFT(ctx, 2, N, C * F, +1)
    img = hip.DeviceBuffer(ctx, 4 * image * image * F)
    graph = hip.Graph(p, s, g, fft, (image, image, 1), True, None, img)
    eager = img.to_host(np.float32).reshape(F, image * image)
    replays = []
    for _ in range(3):
        graph.launch()
        ctx.sync()
        replays.append(img.to_host(np.float32).reshape(F, image * image))
    off = np.concatenate([[0], np.cumsum(sizes)])

    def frame(f):
        d = np.ascontiguousarray(data[:, off[f]:off[f + 1]])
        r32 = orc.grid_f32(tns[f], dcfs[f], d, N, 4.0, lut, nthreads=1)
        r64 = orc.grid_f64(tns[f], dcfs[f], d, N, 4.0, lut, nthreads=1)
        return r32, r64
    worst32 = worst64 = worst_img = 0.0
    with cf.ThreadPoolExecutor(THREADS) as ex:
        for f, (r32, r64) in enumerate(ex.map(frame, range(F))):
            e32, e64, c64 = rel_l2(got[:, f], r32), rel_l2(got[:, f], r64), rel_l2(r32, r64)
            assert e32 <= TOL and e64 <= 2 * c64 + 1e-7, (f, e32, e64, c64)
            worst32, worst64 = max(worst32, e32), max(worst64, e64)
            if f % 8 == 0:          # image parity on every eighth frame
                ri = oracle_image(orc, 2, N, r32, (image, image, 1))
                worst_img = max(worst_img, rel_l2(eager[f], ri))
    for r in replays:
        np.testing.assert_array_equal(r.view(np.int32), eager.view(np.int32))
    for h in (graph, img, fft, g, p, s):
        h.close()
    log(f"[c4] {F} frames ({int(counts.min())}-{int(counts.max())} interleaves x 4096), {M} samples, 12 coils, "
        f"384^2, mode {info['accumulate']}: worst frame rel-L2 vs FP32 {worst32:.2e}, vs FP64 {worst64:.2e}; "
        f"image (every 8th frame) {worst_img:.2e}; 3 graph replays bitwise equal to the eager pass")
    assert worst_img <= TOL


def kdtree_features(tn, N, idx):
    """The oracle's eight features (oracle/ldcf_oracle.py: features) of the
    samples idx of a large trajectory: the same FP32 radius and angles, the
    neighbour counts through a k-d tree over the nearest cells of all samples
    (Chebyshev radius r) instead of the oracle's O(M^2) loop, the shell
    density of the oracle over all samples."""
    import ldcf_oracle as lo
    from scipy.spatial import cKDTree
    dim = tn.shape[1]
    cells = np.clip(np.floor(tn + np.float32(0.5)), 0, N - 1)
    tree = cKDTree(cells)
    f = np.zeros((len(idx), lo.NF), np.float64)
    cf_ = np.float32(0.5) * np.float32(N - 1)
    h = np.float32(0.5) * np.float32(N)

    def radius(t):
        u = (t - cf_) / h
        r2 = (u[:, 0] * u[:, 0] + u[:, 1] * u[:, 1]) + u[:, 2] * u[:, 2] if dim == 3 else u[:, 0] * u[:, 0] + u[:, 1] * u[:, 1]
        return u, np.sqrt(r2)

    u, r = radius(tn[idx])
    f[:, 0] = np.log1p(r.astype(np.float64) * float(h)) / 5.0
    f[:, 1] = np.arctan2(u[:, 1].astype(np.float64), u[:, 0].astype(np.float64)) / np.pi
    if dim == 3:
        with np.errstate(invalid="ignore", divide="ignore"):
            f[:, 2] = np.where(r > 0, np.arccos(np.clip(u[:, 2].astype(np.float64) / r, -1, 1)) / np.pi, 0.5)
    else:
        f[:, 2] = 0.5
    for q, rr in enumerate((1, 2, 4)):
        n = np.asarray(tree.query_ball_point(cells[idx], rr + 0.5, p=np.inf, return_length=True)) - 1
        f[:, 3 + q] = np.log1p(n) / 8.0
    rho_all = (radius(tn)[1] * h).astype(np.float32)
    rho = (r * h).astype(np.float32)
    f[:, 6] = np.minimum(rho_all.max() - rho, np.float32(4.0)).astype(np.float64) * 0.25
    f[:, 7] = lo.shell_density(rho_all, N, dim)[idx]
    return f


# ------------------------------------------------------------------ c5
def test_c5_shipped_network_on_c2_geometry(ctx, orc):
    """The shipped checkpoint on c2's kooshball (32 coils): its DCF against the
    oracle's features + network on a seeded subset of samples, and the image
    against the 10-iteration Pipe-Menon image (NRMSE without a scale fit,
    the measure of this configuration): at most 1% from the network alone and
    with the sampled identity scale (fr_dcf_identity_scale)."""
    import ldcf_oracle as lo
    from fastrecon_amd import ldcf
    N, image = 256, 128
    traj, ramp = syn.kooshball3d(51200, 256)
    tn, _ = hip.normalize_traj(traj, N)
    M = tn.shape[0]
    mlp, meta = ldcf.load(ctx)
    s, p = ldcf.plan_for(ctx, tn, N, 32)
    s.set_coils(coils(32, M, 5))
    mlp.dcf_from_traj(s, N)
    w_ml = s.get_dcf()
    rng = np.random.default_rng(55)
    idx = np.sort(rng.choice(M, 4000, replace=False))
    f = kdtree_features(tn, N, idx)
    params, mu, sigma = mlp.get_params()
    y, _ = lo.forward(params, f.astype(np.float32), emulate=True)
    w_ref = lo.dcf_head(y, f, mu, sigma, mlp.base)
    dw = np.abs(w_ml[idx] / w_ref - 1.0)
    # the image measure
    rc = (image,) * 3
    hip.pipe_menon(p, s, 10, True)
    im_it = ldcf.recon_image(ctx, p, s, N, rc)
    mlp.dcf_from_traj(s, N)
    im_ml = ldcf.recon_image(ctx, p, s, N, rc)
    ldcf.learned_dcf(ctx, mlp, s, N, scale_plan=p)       # + the sampled identity scale
    im_ms = ldcf.recon_image(ctx, p, s, N, rc)
    s.set_dcf(ramp)
    im_rp = ldcf.recon_image(ctx, p, s, N, rc)
    e_ml, e_ms, e_rp = ldcf.nrmse(im_ml, im_it), ldcf.nrmse(im_ms, im_it), ldcf.nrmse(im_rp, im_it)
    es_ml, es_rp = ldcf.nrmse_scaled(im_ml, im_it), ldcf.nrmse_scaled(im_rp, im_it)
    for h in (p, s):
        h.close()
    mlp.close()
    log(f"[c5] shipped network on c2 geometry: DCF vs oracle features+network (4000 samples) median "
        f"{np.median(dw):.2e}, max {dw.max():.2e}; image NRMSE vs iterative: learned {e_ml:.4f} (scaled {es_ml:.4f}, "
        f"sampled identity scale {e_ms:.4f}), ramp {e_rp:.4f} (scaled {es_rp:.4f})")
    # BF16 hidden units: a rounding flip moves y by ~1e-3 (sigma times that in the DCF)
    assert np.median(dw) <= 1e-4 and dw.max() <= 2e-2
    assert e_ml < e_rp and es_ml < es_rp
    assert e_ml <= 0.01 and e_ms <= 0.01


def test_c5_2d_network_on_c1_geometry(ctx, orc):
    """The literal reading of the learned-DCF workload: the 2D radial
    geometry (402 golden-angle spokes x 512, 256^2 image on the 512^2 grid)
    with 32 coils and the shipped 2D network (trained on 2D radial and
    spiral trajectories).  Its DCF against the oracle's features + network
    on seeded samples, and the image against the 10-iteration Pipe-Menon
    image (NRMSE, no scale fit)."""
    import ldcf_oracle as lo
    from fastrecon_amd import ldcf
    N, image = ldcf.N_GRID_2D, ldcf.IMAGE_2D
    traj, ramp = syn.radial2d(402, 512)
    tn, _ = hip.normalize_traj(traj, N)
    M = tn.shape[0]
    mlp, meta = ldcf.load(ctx, ldcf.CHECKPOINT_2D)
    assert meta.get("dim") == 2
    s, p = ldcf.plan_for(ctx, tn, N, 32)
    s.set_coils(coils(32, M, 6))
    mlp.dcf_from_traj(s, N)
    w_ml = s.get_dcf()
    idx = np.sort(np.random.default_rng(56).choice(M, 3000, replace=False))
    f = kdtree_features(tn, N, idx)
    params, mu, sigma = mlp.get_params()
    y, _ = lo.forward(params, f.astype(np.float32), emulate=True)
    w_ref = lo.dcf_head(y, f, mu, sigma, mlp.base)
    dw = np.abs(w_ml[idx] / w_ref - 1.0)
    rc = (image, image, 1)
    hip.pipe_menon(p, s, 10, True)
    im_it = ldcf.recon_image(ctx, p, s, N, rc)
    mlp.dcf_from_traj(s, N)
    im_ml = ldcf.recon_image(ctx, p, s, N, rc)
    hip.dcf_identity_scale(p, s)
    im_mi = ldcf.recon_image(ctx, p, s, N, rc)
    s.set_dcf(ramp)
    im_rp = ldcf.recon_image(ctx, p, s, N, rc)
    e_ml, e_mi, e_rp = ldcf.nrmse(im_ml, im_it), ldcf.nrmse(im_mi, im_it), ldcf.nrmse(im_rp, im_it)
    es_ml, es_rp = ldcf.nrmse_scaled(im_ml, im_it), ldcf.nrmse_scaled(im_rp, im_it)
    for h in (p, s):
        h.close()
    mlp.close()
    log(f"[c5 2D] 2D network on the 402x512 radial geometry, 32 coils: DCF vs oracle features+network (3000 "
        f"samples) median {np.median(dw):.2e}, max {dw.max():.2e}; image NRMSE vs iterative: learned {e_ml:.4f} "
        f"(scaled {es_ml:.4f}, identity-scaled {e_mi:.4f}), ramp {e_rp:.4f} (scaled {es_rp:.4f})")
    assert np.median(dw) <= 1e-4 and dw.max() <= 2e-2
    assert e_ml < e_rp and es_ml < es_rp


def test_c5_shipped_network_phantom_and_psf(ctx):
    """The shipped checkpoint on c2's geometry against the analytic 3D
    ellipsoid phantom (single coil) and the point-spread function (all-ones
    data): NRMSE against the 10-iteration Pipe-Menon image without a scale
    fit, below the analytic ramp DCF's with its least-squares scale; the PSF
    sidelobe level (max outside a 2-pixel ball over the peak) no worse than
    the ramp's."""
    from fastrecon_amd import ldcf
    N, image = 256, 128
    traj, ramp = syn.kooshball3d(51200, 256)
    mlp, _ = ldcf.load(ctx)
    r = ldcf.evaluate_phantom(ctx, mlp, traj, N, (image,) * 3, image, ramp=ramp)
    tn, _ = hip.normalize_traj(traj, N)
    M = tn.shape[0]
    s, p = ldcf.plan_for(ctx, tn, N, 1)
    s.set_coils(np.ones((1, M), np.complex64))
    rc = (image,) * 3
    side = {}
    hip.pipe_menon(p, s, 10, True)
    side["iterative"] = ldcf.psf_sidelobe(ldcf.recon_image(ctx, p, s, N, rc).reshape(rc))
    mlp.dcf_from_traj(s, N)
    side["learned"] = ldcf.psf_sidelobe(ldcf.recon_image(ctx, p, s, N, rc).reshape(rc))
    s.set_dcf(ramp)
    side["ramp"] = ldcf.psf_sidelobe(ldcf.recon_image(ctx, p, s, N, rc).reshape(rc))
    for h in (p, s):
        h.close()
    mlp.close()
    log(f"[c5 phantom] NRMSE vs iterative: learned {r['nrmse_learned']:.4f} (scaled {r['nrmse_learned_scaled']:.4f}, "
        f"identity-scaled {r['nrmse_learned_identity_scaled']:.4f}), ramp {r['nrmse_ramp']:.4f} (scaled "
        f"{r['nrmse_ramp_scaled']:.4f}); PSF sidelobe learned {side['learned']:.4f}, iterative "
        f"{side['iterative']:.4f}, ramp {side['ramp']:.4f}")
    # the learned DCF without any scale fit against the ramp with the best one
    assert r["nrmse_learned"] <= r["nrmse_ramp_scaled"]
    assert r["nrmse_learned"] <= 0.025
    assert side["learned"] <= side["ramp"]
