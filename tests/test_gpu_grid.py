"""GPU parity of the gridding hot path against the oracle, through the C ABI.

Bar: in exact-order mode, bit-identical to the oracle's FP32 restatement of
the reference CPU loop (GridLut.cpp:80-147); in the default mode (heavy units
split by sample range and combined in a fixed order) relative L2 <= 1e-5, the
target FP32 tolerance, and run-to-run bitwise determinism."""
import numpy as np
import pytest

from fastrecon_amd import hip
from fastrecon_amd import synthetic as syn

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module")
def ctx():
    c = hip.Context(0)
    yield c
    c.close()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


EXACT = {"exact_order": 1}


def run_grid(ctx, traj_n, dcf, data, N, lut, kwidth=4.0, tune=EXACT, sample_type=hip.FR_SAMPLES_C64):
    M, dim = traj_n.shape
    C = data.shape[0]
    s = hip.Samples(ctx, dim, M, C, sample_type)
    s.set_traj(traj_n)
    s.set_dcf(dcf)
    s.set_coils(data)
    p = hip.Plan(ctx, s, N, kwidth, lut, tune)
    g = hip.Grids(ctx, dim, N, C)
    hip.grid_execute(p, s, g)
    out = g.download_all()
    info = p.info()
    for h in (g, p, s):
        h.close()
    return out, info


def case(orc, cfg, spokes, readout, coils, seed, image=None):
    c = syn.CONFIGS[cfg]
    image = image or c["image"]
    N = syn.grid_size(image, c["og"])
    traj, dcf = syn.radial2d(spokes, readout) if c["dim"] == 2 else syn.kooshball3d(spokes, readout)
    data = syn.coil_data(coils, traj.shape[0], seed)
    lut, _ = hip.kb_lut(c["kwidth"], c["og"], c["L"])
    tn, _ = hip.normalize_traj(traj, N)
    return tn, dcf, data, N, lut


def check_exact(got, ref, label):
    print(f"[{label}] rel-L2 vs oracle FP32 = {rel_l2(got, ref):.3e}, "
          f"bitwise equal = {np.array_equal(got.view(np.int32), ref.view(np.int32))}", flush=True)
    assert rel_l2(got, ref) <= TOL
    np.testing.assert_array_equal(got.view(np.int32), ref.view(np.int32))


@pytest.mark.parametrize("acc", [0, 4])
def test_radial2d_full_config_bitwise(ctx, orc, acc):
    # the CPU-runnable config: 2D golden-angle, 402 x 512, 8 coils, 512^2 grid
    tn, dcf, data, N, lut = case(orc, "radial2d", 402, 512, 8, seed=1)
    print("radial2d: M=%d N=%d accumulate=%d" % (tn.shape[0], N, acc), flush=True)
    got, info = run_grid(ctx, tn, dcf, data, N, lut, tune=dict(EXACT, accumulate=acc))
    ref = orc.grid_f32(tn, dcf, data, N, 4.0, lut, nthreads=8)
    check_exact(got, ref, "radial2d")
    ref64 = orc.grid_f64(tn, dcf, data, N, 4.0, lut, nthreads=8)
    print(f"  vs FP64 oracle: gpu {rel_l2(got, ref64):.3e}  cpu-fp32 {rel_l2(ref, ref64):.3e}", flush=True)


@pytest.mark.parametrize("coils,acc", [(1, 0), (4, 0), (32, 0), (40, 0), (1, 4), (32, 4), (40, 4)])
def test_kooshball3d_small_bitwise(ctx, orc, coils, acc):
    tn, dcf, data, N, lut = case(orc, "kooshball3d", 600, 64, coils, seed=2, image=32)
    print("kooshball small: M=%d N=%d C=%d accumulate=%d" % (tn.shape[0], N, coils, acc), flush=True)
    got, info = run_grid(ctx, tn, dcf, data, N, lut, tune=dict(EXACT, accumulate=acc))
    ref = orc.grid_f32(tn, dcf, data, N, 4.0, lut, nthreads=8)
    check_exact(got, ref, f"koosh C={coils}")


REGS_OFF = pytest.mark.skip(reason="VGPR accumulation (accumulate=3) gave wrong sums and a GPU memory fault on "
                                   "gfx950 with no established cause; the mode is refused by the plan")


@pytest.mark.parametrize("tune", [dict(accumulate=1), dict(accumulate=2),
                                  pytest.param(dict(accumulate=3), marks=REGS_OFF),
                                  dict(tile_x=8, tile_y=4, tile_z=2), dict(tile_x=4, tile_y=4, tile_z=4),
                                  pytest.param(dict(tile_x=2, tile_y=4, tile_z=4, accumulate=3), marks=REGS_OFF),
                                  dict(tile_x=8, tile_y=2, tile_z=2), dict(tile_x=3, tile_y=3, tile_z=3),
                                  dict(split_taps=64), dict(split_taps=1), dict(split_taps=1, accumulate=2),
                                  dict(accumulate=4), dict(accumulate=4, tile_x=2, tile_y=4, tile_z=4),
                                  dict(accumulate=4, split_taps=64), dict(accumulate=4, tile_x=3, tile_y=3,
                                                                          tile_z=3)])
def test_tile_shape_and_accumulate_mode_invariance(ctx, orc, tune):
    tn, dcf, data, N, lut = case(orc, "kooshball3d", 300, 48, 6, seed=5, image=24)
    got, info = run_grid(ctx, tn, dcf, data, N, lut, tune=dict(tune, exact_order=1))
    assert info["tile_x"] == tune.get("tile_x", info["tile_x"])
    ref = orc.grid_f32(tn, dcf, data, N, 4.0, lut, nthreads=8)
    check_exact(got, ref, f"tune {tune}")


def test_register_mode_is_refused(ctx, orc):
    """The plan refuses VGPR accumulation (accumulate=3) before any launch."""
    tn, dcf, data, N, lut = case(orc, "kooshball3d", 50, 32, 2, seed=5, image=16)
    with pytest.raises(hip.HipError, match="disabled"):
        run_grid(ctx, tn, dcf, data, N, lut, tune=dict(accumulate=3))


def test_edges_integer_centres_and_outside_samples(ctx, orc):
    rng = np.random.default_rng(3)
    N = 40
    pts = rng.uniform(-3.0, N + 2.0, size=(3000, 3)).astype(np.float32)
    pts[:200] = np.round(pts[:200])                  # integer centres: 5-wide boxes
    pts[200:260, 0] = 0.0                            # on the low edge
    pts[260:320, 1] = N - 1.0                        # on the high edge
    dcf = rng.uniform(0, 2, size=len(pts)).astype(np.float32)
    data = syn.coil_data(3, len(pts), 9)
    lut, _ = hip.kb_lut(4.0, 2.0, 512)
    got, _ = run_grid(ctx, pts, dcf, data, N, lut)
    ref = orc.grid_f32(pts, dcf, data, N, 4.0, lut)
    check_exact(got, ref, "edges")


def test_single_sample_and_determinism(ctx, orc):
    lut, _ = hip.kb_lut(4.0, 2.0, 512)
    pt = np.array([[10.3, 11.7, 9.5]], np.float32)
    got, _ = run_grid(ctx, pt, np.ones(1, np.float32), np.ones((1, 1), np.complex64), 24, lut)
    ref = orc.grid_f32(pt, np.ones(1, np.float32), np.ones((1, 1), np.complex64), 24, 4.0, lut)
    check_exact(got, ref, "one sample")
    tn, dcf, data, N, lut = case(orc, "kooshball3d", 400, 64, 8, seed=7, image=32)
    a, _ = run_grid(ctx, tn, dcf, data, N, lut)
    b, _ = run_grid(ctx, tn, dcf, data, N, lut)
    assert np.array_equal(a.view(np.int32), b.view(np.int32))


@pytest.mark.parametrize("acc", [0, 4])
def test_fp16_samples_against_dequantised_oracle(ctx, orc, acc):
    tn, dcf, data, N, lut = case(orc, "kooshball3d_fp16", 500, 96, 16, seed=3, image=48)
    got, _ = run_grid(ctx, tn, dcf, data, N, lut, tune=dict(EXACT, accumulate=acc), sample_type=hip.FR_SAMPLES_C16)
    deq = (data.real.astype(np.float16).astype(np.float32)
           + 1j * data.imag.astype(np.float16).astype(np.float32)).astype(np.complex64)
    ref = orc.grid_f32(tn, dcf, deq, N, 4.0, lut, nthreads=8)
    check_exact(got, ref, "fp16 samples")
    ref64 = orc.grid_f64(tn, dcf, deq, N, 4.0, lut, nthreads=8)
    for c in range(16):
        e = rel_l2(got[c], ref64[c])
        assert e < 1e-6
    print("  per-coil rel-L2 vs FP64 oracle (dequantised inputs):",
          " ".join(f"{rel_l2(got[c], ref64[c]):.2e}" for c in range(16)), flush=True)


def test_update_acquisition(ctx, orc):
    tn, dcf, data, N, lut = case(orc, "radial2d", 40, 64, 2, seed=11, image=32)
    M = tn.shape[0]
    s = hip.Samples(ctx, 2, M, 2)
    s.set_traj(tn); s.set_dcf(dcf); s.set_coils(data)
    new = syn.complex_gaussian(64, 99)
    s.update(1, 64 * 5, new)                      # acquisition 5 of coil 1
    data2 = data.copy(); data2[1, 320:384] = new
    p = hip.Plan(ctx, s, N, 4.0, lut, EXACT)
    g = hip.Grids(ctx, 2, N, 2)
    hip.grid_execute(p, s, g)
    check_exact(g.download_all(), orc.grid_f32(tn, dcf, data2, N, 4.0, lut), "update")


@pytest.mark.parametrize("coils,tune", [(32, {}), (8, {}), (40, {}), (32, dict(split_taps=300)),
                                        (32, dict(split_taps=300, accumulate=2)),
                                        (32, dict(accumulate=4)), (8, dict(accumulate=4)),
                                        (40, dict(accumulate=4, split_taps=300)),
                                        (32, dict(split_taps=2000, tile_x=4, tile_y=4, tile_z=4))])
def test_sliced_mode_parity_and_determinism(ctx, orc, coils, tune):
    """Default mode: heavy units are cut by sample range; the slices are
    summed in a fixed order, so the result is deterministic and within FP32
    rounding of the CPU path, though not in its exact summation order."""
    tn, dcf, data, N, lut = case(orc, "kooshball3d", 900, 64, coils, seed=6, image=32)
    a, info = run_grid(ctx, tn, dcf, data, N, lut, tune=tune)
    b, _ = run_grid(ctx, tn, dcf, data, N, lut, tune=tune)
    assert info["heavy_units"] > 0 and info["items"] > info["streams"], info      # slicing exercised
    assert np.array_equal(a.view(np.int32), b.view(np.int32))
    ref = orc.grid_f32(tn, dcf, data, N, 4.0, lut, nthreads=8)
    ref64 = orc.grid_f64(tn, dcf, data, N, 4.0, lut, nthreads=8)
    e32, e64, c64 = rel_l2(a, ref), rel_l2(a, ref64), rel_l2(ref, ref64)
    print(f"[sliced C={coils} {tune}] heavy={info['heavy_units']} items={info['items']} "
          f"rel-L2 vs CPU-FP32 {e32:.2e}, vs FP64 {e64:.2e} (CPU-FP32 vs FP64 {c64:.2e})", flush=True)
    assert e32 <= TOL
    assert e64 <= 2 * c64 + 1e-7
