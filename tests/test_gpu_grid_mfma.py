"""GPU parity of the matrix-core gridding modes through the C ABI: accumulate
5 (weights from the plan's taps) and 6 (weights computed from the sample
positions, the default).

Both kernels use every tap weight of the reference arithmetic
(tap test sqrtf(d2) < hw, LUT index roundf(dk / hw * (L - 1)), GridLut.cpp:99-
136) and multiplies it with dcf * sample (GridLut.cpp:109) like the CPU loop;
only the order of the sum over samples differs (four products per matrix
instruction).  So:

* where every grid cell receives at most one sample (samples at least four
  cells apart in some axis), each cell is one product and the result must be
  bit-identical to the oracle -- this pins the tap set and every LUT index,
  including integer and half-integer centres and taps at distance exactly hw;
* in general the result agrees with the oracle's FP32 loop to relative L2
  <= 1e-5 and is no further from the FP64 oracle than twice the FP32 loop is,
  and it is bitwise reproducible run to run."""
import numpy as np
import pytest

from fastrecon_amd import hip
from fastrecon_amd import synthetic as syn

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(params=[5, 6], ids=["taps", "positions"])
def mode(request):
    return request.param


def cells_of(tune, dim):
    d = dict(tune or {})
    bx, by = d.get("tile_x", 4), d.get("tile_y", 4 if dim == 2 else 2)
    bz = d.get("tile_z", 2) if dim == 3 else 1
    return bx * by * bz


def skip_big_units(mode, tune, dim):
    if mode == 6 and cells_of(tune, dim) > 16:
        pytest.skip("the position-weight mode takes unit boxes of at most 16 cells")


@pytest.fixture(scope="module")
def ctx():
    c = hip.Context(0)
    yield c
    c.close()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


def run_grid(ctx, traj_n, dcf, data, N, lut, kwidth=4.0, tune=None, sample_type=hip.FR_SAMPLES_C64, mode=5):
    M, dim = traj_n.shape
    C = data.shape[0]
    s = hip.Samples(ctx, dim, M, C, sample_type)
    s.set_traj(traj_n)
    s.set_dcf(dcf)
    s.set_coils(data)
    p = hip.Plan(ctx, s, N, kwidth, lut, dict(tune or {}, accumulate=mode))
    g = hip.Grids(ctx, dim, N, C)
    hip.grid_execute(p, s, g)
    out = g.download_all()
    info = p.info()
    for h in (g, p, s):
        h.close()
    return out, info


def case(cfg, spokes, readout, coils, seed, image=None):
    c = syn.CONFIGS[cfg]
    image = image or c["image"]
    N = syn.grid_size(image, c["og"])
    traj, dcf = syn.radial2d(spokes, readout) if c["dim"] == 2 else syn.kooshball3d(spokes, readout)
    data = syn.coil_data(coils, traj.shape[0], seed)
    lut, _ = hip.kb_lut(c["kwidth"], c["og"], c["L"])
    tn, _ = hip.normalize_traj(traj, N)
    return tn, dcf, data, N, lut


def sparse_points(dim, N, spacing, rng, jitter=1.0):
    """Jittered lattice: any two samples are >= spacing - jitter cells apart
    in some axis (more than the kernel width), so no grid cell is tapped by
    two samples.  A quarter of the samples sit
    on integer or half-integer centres (5-wide boxes, taps at distance exactly
    hw = 2 that must be rejected, LUT indices on rounding ties)."""
    axes = [np.arange(2.0, N - 3.0, spacing)] * dim
    grid = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, dim)
    jit = rng.uniform(0.0, jitter, size=grid.shape)
    pts = grid + jit
    n = len(pts)
    q = n // 4
    pts[:q] = np.floor(pts[:q])                        # integer centres
    pts[q:2 * q] = np.floor(pts[q:2 * q]) + 0.5        # half-integer centres
    return pts.astype(np.float32)


@pytest.mark.parametrize("dim,coils,tune", [
    (3, 32, {}), (3, 8, {}), (3, 12, {}), (3, 1, {}), (3, 40, {}),
    (3, 32, dict(tile_x=4, tile_y=4, tile_z=2)), (3, 32, dict(tile_x=4, tile_y=4, tile_z=4)),
    (3, 16, dict(tile_x=3, tile_y=3, tile_z=3)), (3, 32, dict(tile_x=8, tile_y=2, tile_z=1)),
    (3, 32, dict(tile_x=2, tile_y=2, tile_z=4)), (3, 8, dict(tile_x=4, tile_y=4, tile_z=1)),
    (3, 32, dict(tile_writes=1)), (3, 16, dict(tile_writes=1, stores=1)),
    (2, 8, {}), (2, 12, dict(tile_x=8, tile_y=4)), (2, 32, dict(tile_x=8, tile_y=8)), (2, 16, dict(tile_x=8, tile_y=2)),
    (2, 8, dict(tile_writes=1)), (3, 32, dict(tile_writes=2)), (3, 12, dict(tile_writes=3)),
    (3, 1, dict(tile_writes=3)), (3, 33, dict(tile_writes=3)), (2, 12, dict(tile_writes=3))])
def test_sparse_samples_bitwise(ctx, orc, mode, dim, coils, tune):
    skip_big_units(mode, tune, dim)
    rng = np.random.default_rng(11 + dim + coils)
    N = 48 if dim == 3 else 96
    pts = sparse_points(dim, N, 5.0, rng)
    dcf = rng.uniform(0.1, 2.0, size=len(pts)).astype(np.float32)
    data = syn.coil_data(coils, len(pts), 21)
    lut, _ = hip.kb_lut(4.0, 2.0, 512)
    got, info = run_grid(ctx, pts, dcf, data, N, lut, tune=tune, mode=mode)
    ref = orc.grid_f32(pts, dcf, data, N, 4.0, lut, nthreads=8)
    same = np.array_equal(got.view(np.int32), ref.view(np.int32))
    print(f"[sparse mode {mode} {dim}D C={coils} {tune}] M={len(pts)} tile={info['tile_x']}x{info['tile_y']}x{info['tile_z']} "
          f"rel-L2 {rel_l2(got, ref):.2e} bitwise {same}", flush=True)
    np.testing.assert_array_equal(got.view(np.int32), ref.view(np.int32))


@pytest.mark.parametrize("coils,tune", [(32, {}), (1, {}), (2, {}), (8, {}), (12, {}), (16, {}), (40, {}), (33, {}),
                                        (34, {}),
                                        (32, dict(split_taps=300)), (32, dict(tile_x=4, tile_y=4, tile_z=2)),
                                        (32, dict(tile_x=4, tile_y=4, tile_z=4, split_taps=2000)),
                                        (12, dict(tile_x=3, tile_y=3, tile_z=3)),
                                        (32, dict(split_taps=16)), (32, dict(items_per_wave=3, xcd_run=40)),
                                        (32, dict(tile_writes=1)), (12, dict(tile_writes=1, split_taps=64)),
                                        (32, dict(tile_writes=1, stores=1, waves_per_simd=3)),
                                        (8, dict(tile_writes=1, tile_x=4, tile_y=4, tile_z=1)),
                                        (32, dict(stores=1, waves_per_simd=3, items_per_wave=2)),
                                        (32, dict(tile_x=8, tile_y=2, tile_z=1, items_per_wave=2)),
                                        (16, dict(tile_x=2, tile_y=2, tile_z=4, split_taps=100)),
                                        (32, dict(tile_writes=2)), (16, dict(tile_writes=3, split_taps=64)),
                                        (5, dict(tile_writes=3)), (33, dict(tile_writes=3, items_per_wave=2))])
def test_kooshball_parity_and_determinism(ctx, orc, mode, coils, tune):
    skip_big_units(mode, tune, 3)
    tn, dcf, data, N, lut = case("kooshball3d", 900, 64, coils, seed=6, image=32)
    a, info = run_grid(ctx, tn, dcf, data, N, lut, tune=tune, mode=mode)
    b, _ = run_grid(ctx, tn, dcf, data, N, lut, tune=tune, mode=mode)
    assert np.array_equal(a.view(np.int32), b.view(np.int32))
    ref = orc.grid_f32(tn, dcf, data, N, 4.0, lut, nthreads=8)
    ref64 = orc.grid_f64(tn, dcf, data, N, 4.0, lut, nthreads=8)
    e32, e64, c64 = rel_l2(a, ref), rel_l2(a, ref64), rel_l2(ref, ref64)
    print(f"[mode {mode} C={coils} {tune}] heavy={info['heavy_units']} items={info['items']} "
          f"rel-L2 vs CPU-FP32 {e32:.2e}, vs FP64 {e64:.2e} (CPU-FP32 vs FP64 {c64:.2e})", flush=True)
    assert e32 <= TOL
    assert e64 <= 2 * c64 + 1e-7


def test_radial2d_full_config(ctx, orc, mode):
    """Config-1 geometry: 402 golden-angle spokes x 512, 8 coils, 512^2 grid."""
    tn, dcf, data, N, lut = case("radial2d", 402, 512, 8, seed=1)
    got, info = run_grid(ctx, tn, dcf, data, N, lut, mode=mode)
    ref = orc.grid_f32(tn, dcf, data, N, 4.0, lut, nthreads=8)
    ref64 = orc.grid_f64(tn, dcf, data, N, 4.0, lut, nthreads=8)
    e32, e64, c64 = rel_l2(got, ref), rel_l2(got, ref64), rel_l2(ref, ref64)
    print(f"[mode {mode} radial2d] M={tn.shape[0]} N={N} items={info['items']} rel-L2 vs CPU-FP32 {e32:.2e}, "
          f"vs FP64 {e64:.2e} (CPU-FP32 vs FP64 {c64:.2e})", flush=True)
    assert e32 <= TOL
    assert e64 <= 2 * c64 + 1e-7


def test_edges_integer_centres_and_outside_samples(ctx, orc, mode):
    rng = np.random.default_rng(3)
    N = 40
    pts = rng.uniform(-3.0, N + 2.0, size=(3000, 3)).astype(np.float32)
    pts[:200] = np.round(pts[:200])
    pts[200:260, 0] = 0.0
    pts[260:320, 1] = N - 1.0
    dcf = rng.uniform(0, 2, size=len(pts)).astype(np.float32)
    data = syn.coil_data(3, len(pts), 9)
    lut, _ = hip.kb_lut(4.0, 2.0, 512)
    got, _ = run_grid(ctx, pts, dcf, data, N, lut, mode=mode)
    ref = orc.grid_f32(pts, dcf, data, N, 4.0, lut)
    print(f"[mode {mode} edges] rel-L2 {rel_l2(got, ref):.2e}", flush=True)
    assert rel_l2(got, ref) <= TOL


@pytest.mark.parametrize("kw,og,L", [(3.0, 2.0, 512), (5.0, 1.5, 256), (4.0, 1.5, 1024)])
def test_other_kernel_widths_and_lut_lengths(ctx, orc, mode, kw, og, L):
    """hw that is not a power of two (1.5, 2.5) and other LUT lengths: the
    exact index thresholds are built from the reference arithmetic."""
    rng = np.random.default_rng(int(kw * 10 + L))
    N = 48
    pts = sparse_points(3, N, 7.0, rng)
    dcf = rng.uniform(0.1, 2.0, size=len(pts)).astype(np.float32)
    data = syn.coil_data(8, len(pts), 4)
    lut, _ = hip.kb_lut(kw, og, L)
    got, _ = run_grid(ctx, pts, dcf, data, N, lut, kwidth=kw, mode=mode)
    ref = orc.grid_f32(pts, dcf, data, N, kw, lut)
    np.testing.assert_array_equal(got.view(np.int32), ref.view(np.int32))


def test_fp16_samples(ctx, orc, mode):
    tn, dcf, data, N, lut = case("kooshball3d_fp16", 500, 96, 16, seed=3, image=48)
    got, _ = run_grid(ctx, tn, dcf, data, N, lut, sample_type=hip.FR_SAMPLES_C16, mode=mode)
    deq = (data.real.astype(np.float16).astype(np.float32)
           + 1j * data.imag.astype(np.float16).astype(np.float32)).astype(np.complex64)
    ref = orc.grid_f32(tn, dcf, deq, N, 4.0, lut, nthreads=8)
    ref64 = orc.grid_f64(tn, dcf, deq, N, 4.0, lut, nthreads=8)
    e = [rel_l2(got[c], ref64[c]) for c in range(16)]
    print(f"[mode {mode} fp16] rel-L2 vs CPU-FP32 {rel_l2(got, ref):.2e}; per coil vs FP64:",
          " ".join(f"{x:.2e}" for x in e), flush=True)
    assert rel_l2(got, ref) <= TOL
    assert max(e) < 1e-6


def test_exact_order_is_refused(ctx, orc, mode):
    tn, dcf, data, N, lut = case("kooshball3d", 50, 32, 2, seed=5, image=16)
    with pytest.raises(hip.HipError, match="exact_order"):
        run_grid(ctx, tn, dcf, data, N, lut, tune=dict(exact_order=1), mode=mode)


def test_position_mode_refuses_big_units_and_long_luts(ctx, orc):
    tn, dcf, data, N, lut = case("kooshball3d", 50, 32, 2, seed=5, image=16)
    with pytest.raises(hip.HipError, match="16 cells"):
        run_grid(ctx, tn, dcf, data, N, lut, tune=dict(tile_x=8, tile_y=2, tile_z=2), mode=6)
    lut4k, _ = hip.kb_lut(4.0, 2.0, 4096)
    with pytest.raises(hip.HipError, match="2048"):
        run_grid(ctx, tn, dcf, data, N, lut4k, mode=6)


def test_default_plan_is_position_mode(ctx, orc):
    tn, dcf, data, N, lut = case("kooshball3d", 60, 32, 4, seed=5, image=16)
    s = hip.Samples(ctx, 3, tn.shape[0], 4)
    s.set_traj(tn)
    p = hip.Plan(ctx, s, N, 4.0, lut)
    assert p.info()["accumulate"] == 6
    assert (p.info()["tile_x"], p.info()["tile_y"], p.info()["tile_z"]) == (4, 2, 2)
    p8 = hip.Plan(ctx, s, N, 4.0, lut, dict(tile_x=8, tile_y=2, tile_z=2))
    assert p8.info()["accumulate"] == 5
    for h in (p8, p, s):
        h.close()


def test_nan_samples_stay_in_the_units_they_touch(ctx, orc, mode):
    """Documented deviation of the matrix-core modes: a NaN sample value
    reaches every cell of the units its taps touch (0 * NaN in the matrix
    product), but no cell outside them, and other coils are unaffected."""
    rng = np.random.default_rng(8)
    N = 32
    pts = sparse_points(3, N, 9.0, rng)
    dcf = np.ones(len(pts), np.float32)
    data = syn.coil_data(2, len(pts), 4)
    data[0, 5] = np.nan
    lut, _ = hip.kb_lut(4.0, 2.0, 512)
    got, info = run_grid(ctx, pts, dcf, data, N, lut, mode=mode)
    ref = orc.grid_f32(pts, dcf, data, N, 4.0, lut)
    bad = ~np.isfinite(got[0])
    refbad = ~np.isfinite(ref[0])
    assert refbad.sum() > 0 and np.all(bad[refbad])
    # every non-finite cell lies within 4 cells of the NaN sample (its
    # footprint widened to whole units)
    cz, cy, cx = np.unravel_index(np.nonzero(bad)[0], (N, N, N))
    d = np.maximum.reduce([np.abs(cx - pts[5, 0]), np.abs(cy - pts[5, 1]), np.abs(cz - pts[5, 2])])
    assert d.max() < 2 + 8
    fin = ~bad
    assert rel_l2(got[0][fin], ref[0][fin]) <= TOL
    assert rel_l2(got[1], ref[1]) <= TOL


def test_dcf_changes_reach_the_position_plan(ctx, orc):
    """The position-weight plan keeps a copy of the DCF with its positions;
    every way a DCF changes reaches the next pass: a host upload, a device
    upload, another samples handle on the same plan, and back."""
    C = 8
    tn, dcf, data, N, lut = case("kooshball3d", 600, 48, C, seed=9, image=24)
    M = tn.shape[0]
    rng = np.random.default_rng(71)
    s = hip.Samples(ctx, 3, M, C)
    s.set_traj(tn)
    s.set_dcf(dcf)
    s.set_coils(data)
    p = hip.Plan(ctx, s, N, 4.0, lut)
    assert p.info()["accumulate"] == 6
    g = hip.Grids(ctx, 3, N, C)

    def check(samples, w, what):
        hip.grid_execute(p, samples, g)
        got = g.download_all()
        ref = orc.grid_f32(tn, w, data, N, 4.0, lut, nthreads=8)
        e = rel_l2(got, ref)
        print(f"[dcf change: {what}] rel-L2 {e:.2e}", flush=True)
        assert e <= TOL

    check(s, dcf, "initial")
    w2 = rng.uniform(0.1, 2.0, M).astype(np.float32)
    s.set_dcf(w2)
    check(s, w2, "host upload")
    w3 = rng.uniform(0.1, 2.0, M).astype(np.float32)
    buf = hip.DeviceBuffer(ctx, 4 * M)
    buf.from_host(w3)
    s.set_dcf_device(buf.ptr)
    check(s, w3, "device upload")
    s2 = hip.Samples(ctx, 3, M, C)
    s2.set_traj(tn)
    w4 = rng.uniform(0.1, 2.0, M).astype(np.float32)
    s2.set_dcf(w4)
    s2.set_coils(data)
    check(s2, w4, "second handle")
    check(s, w3, "first handle again")
    for h in (buf, s2, g, p, s):
        h.close()
