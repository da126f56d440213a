"""Batched uneven 2D frames in one launch, captured HIP-graph replays and
the autotuner, through the C ABI.

Bars: batched gridding of every frame is bit-identical to gridding that
frame on its own (exact order), which is bit-identical to the oracle; the
batched images are within rel-L2 1e-5 of the oracle per frame; every graph
replay equals the eager pass bitwise."""
import numpy as np
import pytest

from fastrecon_amd import hip
from fastrecon_amd import synthetic as syn

pytestmark = pytest.mark.gpu
EXACT = {"exact_order": 1}


@pytest.fixture(scope="module")
def ctx():
    c = hip.Context(0)
    yield c
    c.close()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def frames_case(F=6, samples=512, C=3, image=32, seed=6):
    N = syn.grid_size(image, 2.0)
    frames, counts = syn.spiral_frames(F, samples, seed=seed, lo=2, hi=9)
    lut, _ = hip.kb_lut(4.0, 2.0, 512)
    tns, dcfs, datas = [], [], []
    for f, (k, dcf) in enumerate(frames):
        tn, _ = hip.normalize_traj(k, N)          # each frame with its own bounds
        tns.append(tn)
        dcfs.append(dcf)
        datas.append(syn.coil_data(C, tn.shape[0], 100 + f))
    return N, lut, tns, dcfs, datas, image


def batched_samples(ctx, tns, dcfs, datas, C):
    sizes = np.array([t.shape[0] for t in tns], np.int64)
    M = int(sizes.sum())
    s = hip.Samples(ctx, 2, M, C, frame_samples=sizes)
    s.set_traj(np.concatenate(tns))
    s.set_dcf(np.concatenate(dcfs))
    s.set_coils(np.concatenate(datas, axis=1))
    return s, sizes


@pytest.mark.parametrize("tune", [EXACT, {}])
def test_batched_frames_equal_single_frames(ctx, orc, tune):
    C = 3
    N, lut, tns, dcfs, datas, image = frames_case(C=C)
    s, sizes = batched_samples(ctx, tns, dcfs, datas, C)
    F = len(tns)
    p = hip.Plan(ctx, s, N, 4.0, lut, tune)
    g = hip.Grids(ctx, 2, N, C, frames=F)
    hip.grid_execute(p, s, g)
    got = g.download_all().reshape(C, F, N * N)
    print(f"[frames {tune}] F={F} sizes={sizes.tolist()} items={p.info()['items']}", flush=True)
    for f in range(F):
        ref = orc.grid_f32(tns[f], dcfs[f], datas[f], N, 4.0, lut, nthreads=4)
        if tune == EXACT:
            s1 = hip.Samples(ctx, 2, tns[f].shape[0], C)
            s1.set_traj(tns[f]); s1.set_dcf(dcfs[f]); s1.set_coils(datas[f])
            p1 = hip.Plan(ctx, s1, N, 4.0, lut, EXACT)
            g1 = hip.Grids(ctx, 2, N, C)
            hip.grid_execute(p1, s1, g1)
            one = g1.download_all()
            assert np.array_equal(got[:, f].view(np.int32), one.view(np.int32))
            assert np.array_equal(got[:, f].view(np.int32), ref.view(np.int32))
        else:
            assert rel_l2(got[:, f], ref) <= 1e-5


@pytest.mark.parametrize("image", [32, 128])
def test_batched_recon_and_graph_replays(ctx, orc, image):
    """image 32 (N = 64): rocFFT + the SOS epilogue; image 128 (N = 256):
    the cropped transforms with the coil sum in the last pass."""
    C = 3
    N, lut, tns, dcfs, datas, image = frames_case(C=C, F=5, image=image)
    s, sizes = batched_samples(ctx, tns, dcfs, datas, C)
    F = len(tns)
    p = hip.Plan(ctx, s, N, 4.0, lut)
    g = hip.Grids(ctx, 2, N, C, frames=F)
    fft = hip.FFT(ctx, 2, N, C * F, +1)
    out = hip.DeviceBuffer(ctx, 4 * image * image * F)
    graph = hip.Graph(p, s, g, fft, (image, image, 1), True, None, out)   # runs one eager pass first
    eager = out.to_host(np.float32).reshape(F, image * image)
    for f in range(F):
        ref = orc.recon_f32(tns[f], dcfs[f], datas[f], N, 4.0, lut, (image, image, 1), nthreads=4)
        e = rel_l2(eager[f], ref)
        assert e <= 1e-5, (f, e)
    # replays, with eager work and synchronisations in between
    for r in range(4):
        out.from_host(np.zeros(image * image * F, np.float32))
        graph.launch()
        ctx.sync()
        rep = out.to_host(np.float32).reshape(F, image * image)
        assert np.array_equal(rep, eager), f"replay {r} differs"
        hip.grid_execute(p, s, g)            # unrelated eager pass on the same stream
        ctx.sync()
    # new data in place: the replay grids it
    datas2 = [syn.coil_data(C, t.shape[0], 300 + f) for f, t in enumerate(tns)]
    s.set_coils(np.concatenate(datas2, axis=1))
    graph.launch()
    ctx.sync()
    rep = out.to_host(np.float32).reshape(F, image * image)
    for f in range(F):
        ref = orc.recon_f32(tns[f], dcfs[f], datas2[f], N, 4.0, lut, (image, image, 1), nthreads=4)
        assert rel_l2(rep[f], ref) <= 1e-5
    # a new DCF written on the device after capture: the replay grids with it
    dcfs3 = [(d * np.float32(0.5) + np.float32(0.25)).astype(np.float32) for d in dcfs]
    dbuf = hip.DeviceBuffer(ctx, 4 * sum(len(d) for d in dcfs3))
    dbuf.from_host(np.concatenate(dcfs3))
    s.set_dcf_device(dbuf.ptr)
    graph.launch()
    ctx.sync()
    rep = out.to_host(np.float32).reshape(F, image * image)
    for f in range(F):
        ref = orc.recon_f32(tns[f], dcfs3[f], datas2[f], N, 4.0, lut, (image, image, 1), nthreads=4)
        assert rel_l2(rep[f], ref) <= 1e-5
    dbuf.close()
    graph.close()


def test_pipe_menon_on_batched_frames(ctx):
    """Pipe-Menon on a batched series equals Pipe-Menon per frame, bitwise."""
    N, lut, tns, dcfs, datas, image = frames_case(C=1, F=4)
    s, sizes = batched_samples(ctx, tns, dcfs, datas, 1)
    p = hip.Plan(ctx, s, N, 4.0, lut, EXACT)
    hip.pipe_menon(p, s, 3, True)
    wb = s.get_dcf()
    o = 0
    for f, tn in enumerate(tns):
        s1 = hip.Samples(ctx, 2, tn.shape[0], 1)
        s1.set_traj(tn); s1.set_dcf(dcfs[f])
        p1 = hip.Plan(ctx, s1, N, 4.0, lut, EXACT)
        hip.pipe_menon(p1, s1, 3, True)
        assert np.array_equal(wb[o:o + tn.shape[0]], s1.get_dcf())
        o += tn.shape[0]


def test_pipe_menon_on_batched_frames_sliced_plan(ctx):
    """Default (sliced) plans run the density gridding on the matrix cores,
    whose sums group records four at a time: batched and per-frame results
    then agree to FP32 rounding rather than bitwise."""
    N, lut, tns, dcfs, datas, image = frames_case(C=1, F=4)
    s, sizes = batched_samples(ctx, tns, dcfs, datas, 1)
    p = hip.Plan(ctx, s, N, 4.0, lut)
    hip.pipe_menon(p, s, 3, True)
    wb = s.get_dcf()
    o = 0
    for f, tn in enumerate(tns):
        s1 = hip.Samples(ctx, 2, tn.shape[0], 1)
        s1.set_traj(tn); s1.set_dcf(dcfs[f])
        p1 = hip.Plan(ctx, s1, N, 4.0, lut)
        hip.pipe_menon(p1, s1, 3, True)
        w1 = s1.get_dcf()
        e = np.linalg.norm(wb[o:o + tn.shape[0]] - w1) / np.linalg.norm(w1)
        assert e < 1e-6, e
        o += tn.shape[0]


def test_autotune_picks_a_working_candidate(ctx, orc):
    C = 3
    N, lut, tns, dcfs, datas, image = frames_case(C=C, F=4)
    s, sizes = batched_samples(ctx, tns, dcfs, datas, C)
    best, ms, cand = hip.autotune(ctx, s, N, 4.0, lut, reps=2)
    print(f"[autotune] best {best} {ms:.3f} ms; candidates {cand.tolist()}", flush=True)
    assert ms > 0 and (cand > 0).sum() >= 4
    assert ms == pytest.approx(cand[cand > 0].min())
    # the winner grids correctly; sequential modes also in exact CPU order
    # (the matrix-core mode refuses exact_order: it sums four products at a time)
    exact = best["accumulate"] not in (5, 6)
    p = hip.Plan(ctx, s, N, 4.0, lut, dict(best, exact_order=1) if exact else best)
    g = hip.Grids(ctx, 2, N, C, frames=len(tns))
    hip.grid_execute(p, s, g)
    got = g.download_all().reshape(C, len(tns), N * N)
    ref = orc.grid_f32(tns[0], dcfs[0], datas[0], N, 4.0, lut, nthreads=4)
    if exact:
        assert np.array_equal(got[:, 0].view(np.int32), ref.view(np.int32))
    else:
        e = np.linalg.norm(got[:, 0] - ref) / np.linalg.norm(ref)
        assert e <= 1e-5, e


def test_autotune_search_and_cache(ctx, orc, tmp_path):
    """fr_autotune_search on a small 3D kooshball: the staged search times
    the default and its candidates, its winner grids correctly, and with a
    cache file a second call returns the stored tune without timing."""
    import time
    C, N = 4, 48
    traj, dcf = syn.kooshball3d(1500, 48)
    tn, _ = hip.normalize_traj(traj, N)
    M = tn.shape[0]
    data = syn.coil_data(C, M, 21)
    s = hip.Samples(ctx, 3, M, C)
    s.set_traj(tn)
    s.set_dcf(dcf)
    s.set_coils(data)
    lut, _ = hip.kb_lut(4.0, 2.0, 512)
    path = str(tmp_path / "tune.txt")
    t0 = time.perf_counter()
    best, best_ms, def_ms, timed = hip.autotune_search(ctx, s, N, 4.0, lut, reps=2, cache_path=path)
    t_search = time.perf_counter() - t0
    t0 = time.perf_counter()
    best2, best_ms2, def_ms2, timed2 = hip.autotune_search(ctx, s, N, 4.0, lut, reps=2, cache_path=path)
    t_hit = time.perf_counter() - t0
    print(f"[autotune search] {timed} timed in {t_search:.2f} s: best {best} {best_ms:.4f} ms, default "
          f"{def_ms:.4f} ms; cache hit in {t_hit * 1e3:.2f} ms", flush=True)
    assert timed >= 9 and def_ms > 0 and 0 < best_ms <= 1.1 * def_ms
    assert timed2 == 0 and best2 == best and best_ms2 == pytest.approx(best_ms, rel=1e-5)
    assert def_ms2 == pytest.approx(def_ms, rel=1e-5) and t_hit < 0.1 * t_search
    # the winner grids correctly
    p = hip.Plan(ctx, s, N, 4.0, lut, best)
    g = hip.Grids(ctx, 3, N, C)
    hip.grid_execute(p, s, g)
    got = g.download_all()
    ref = orc.grid_f64(tn, dcf, data, N, 4.0, lut, nthreads=8)
    e = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert e <= 1e-6, e
    for h in (g, p, s):
        h.close()
