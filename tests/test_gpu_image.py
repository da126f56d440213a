"""GPU parity of the image stage and the net-new DCF machinery, through the
C ABI: batched rocFFT (sign +1, unnormalised), fftshift, crop + SOS, the
deapodisation map, degridding and Pipe-Menon, against the oracle.

Bar: relative L2 <= 1e-5 (the target FP32 tolerance) for every image; the
oracle's FFT is FP64, so these numbers include the GPU's FP32 FFT rounding."""
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
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def setup(ctx, dim, spokes, readout, coils, image, seed):
    N = syn.grid_size(image, 2.0)
    traj, dcf = syn.radial2d(spokes, readout) if dim == 2 else syn.kooshball3d(spokes, readout)
    lut, _ = hip.kb_lut(4.0, 2.0, 512)
    tn, _ = hip.normalize_traj(traj, N)
    data = syn.coil_data(coils, tn.shape[0], seed)
    s = hip.Samples(ctx, dim, tn.shape[0], coils)
    s.set_traj(tn)
    s.set_dcf(dcf)
    s.set_coils(data)
    p = hip.Plan(ctx, s, N, 4.0, lut)
    g = hip.Grids(ctx, dim, N, coils)
    return tn, dcf, data, N, lut, s, p, g


@pytest.mark.parametrize("dim,spokes,readout,coils,image", [(2, 402, 512, 8, 256), (3, 800, 64, 4, 32)])
def test_recon_matches_oracle(ctx, orc, dim, spokes, readout, coils, image):
    tn, dcf, data, N, lut, s, p, g = setup(ctx, dim, spokes, readout, coils, image, 21)
    rc = (image, image, image if dim == 3 else 1)
    hip.grid_execute(p, s, g)
    f = hip.FFT(ctx, dim, N, coils, +1)
    f(g)
    img = hip.sos(ctx, g, rc, shift=True).to_host(np.float32)
    ref = orc.recon_f32(tn, dcf, data, N, 4.0, lut, rc, nthreads=8)
    e = rel_l2(img, ref)
    print(f"[recon dim={dim}] image rel-L2 vs oracle = {e:.2e}", flush=True)
    assert e <= TOL
    # explicit in-place fftshift then unshifted SOS gives the same image
    hip.fftshift(ctx, g)
    img2 = hip.sos(ctx, g, rc, shift=False).to_host(np.float32)
    assert np.array_equal(img, img2)


def test_fft_alone_against_numpy(ctx):
    rng = np.random.default_rng(4)
    for dim, N, B in [(2, 384, 3), (3, 48, 2)]:
        x = (rng.standard_normal((B, N ** dim)) + 1j * rng.standard_normal((B, N ** dim))).astype(np.complex64)
        g = hip.Grids(ctx, dim, N, B)
        for c in range(B):
            g.upload(c, x[c])
        f = hip.FFT(ctx, dim, N, B, +1)
        f(g)
        got = g.download_all().reshape((B,) + (N,) * dim)
        ref = np.fft.ifftn(x.reshape((B,) + (N,) * dim).astype(np.complex128), axes=tuple(range(1, dim + 1))) * N ** dim
        e = rel_l2(got, ref)
        print(f"[fft dim={dim} N={N}] rel-L2 vs numpy FP64 = {e:.2e}", flush=True)
        assert e <= TOL


def test_deapod_map_matches_oracle(ctx, orc):
    for dim, N, rc in [(2, 128, (64, 64, 1)), (3, 48, (24, 24, 24))]:
        lut, _ = hip.kb_lut(4.0, 2.0, 512)
        centre = np.full(dim, (N - 1) / 2.0, np.float32)
        got = hip.deapod_map(ctx, dim, N, 4.0, lut, centre, rc).to_host(np.float32)
        grid = orc.grid_f32(centre[None, :], np.ones(1, np.float32), np.ones((1, 1), np.complex64), N, 4.0, lut)
        img = orc.fftshift(dim, N, orc.fft_c2c(dim, N, grid, +1))
        ref = orc.sos(dim, N, img, rc)
        e = rel_l2(got, ref)
        print(f"[deapod dim={dim}] rel-L2 = {e:.2e}, min/max = {got.min():.3f}/{got.max():.3f}", flush=True)
        assert e <= TOL and got.min() > 0


def test_degrid_is_adjoint_and_matches_oracle(ctx, orc):
    tn, dcf, data, N, lut, s, p, g = setup(ctx, 3, 300, 48, 3, 24, 5)
    rng = np.random.default_rng(8)
    y = (rng.standard_normal((3, N ** 3)) + 1j * rng.standard_normal((3, N ** 3))).astype(np.complex64)
    for c in range(3):
        g.upload(c, y[c])
    out = hip.DeviceBuffer(ctx, 8 * tn.shape[0] * 3)
    hip.degrid_execute(p, s, g, out)
    got = out.to_host(np.complex64).reshape(tn.shape[0], 3).T
    ref = orc.degrid_f64(tn, y.astype(np.complex128), N, 4.0, lut)
    e = rel_l2(got, ref)
    print(f"[degrid] rel-L2 vs FP64 oracle = {e:.2e}", flush=True)
    assert e <= TOL
    # adjoint identity against the GPU's own gridding with unit DCF
    s.set_dcf(np.ones(tn.shape[0], np.float32))
    hip.grid_execute(p, s, g)
    Gx = g.download_all().astype(np.complex128)
    lhs = np.vdot(y.astype(np.complex128), Gx)
    rhs = np.vdot(got.astype(np.complex128), data.astype(np.complex128))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


@pytest.mark.parametrize("dim", [2, 3])
def test_degrid_bitwise_against_fp32_oracle(ctx, orc, dim):
    """Degridding takes the LUT index from a fast square root corrected by the
    plan's exact thresholds; taps, indices and summation order are the
    reference's, so it equals the FP32 oracle bit for bit."""
    if dim == 2:
        tn, dcf, data, N, lut, s, p, g = setup(ctx, 2, 80, 96, 3, 48, 4)
    else:
        tn, dcf, data, N, lut, s, p, g = setup(ctx, 3, 300, 48, 3, 24, 5)
    rng = np.random.default_rng(18)
    y = (rng.standard_normal((3, N ** dim)) + 1j * rng.standard_normal((3, N ** dim))).astype(np.complex64)
    for c in range(3):
        g.upload(c, y[c])
    out = hip.DeviceBuffer(ctx, 8 * tn.shape[0] * 3)
    hip.degrid_execute(p, s, g, out)
    got = np.ascontiguousarray(out.to_host(np.complex64).reshape(tn.shape[0], 3).T)
    ref = orc.degrid_f32(tn, y, N, 4.0, lut)
    print(f"[degrid {dim}D] bitwise = {np.array_equal(got.view(np.int32), ref.view(np.int32))}", flush=True)
    np.testing.assert_array_equal(got.view(np.int32), ref.view(np.int32))


@pytest.mark.parametrize("dim", [2, 3])
def test_pipe_menon_matches_oracle(ctx, orc, dim):
    if dim == 2:
        tn, dcf, data, N, lut, s, p, g = setup(ctx, 2, 60, 128, 1, 64, 3)
    else:
        tn, dcf, data, N, lut, s, p, g = setup(ctx, 3, 300, 32, 1, 16, 3)
    hip.pipe_menon(p, s, iters=10, init_ones=True)
    got = s.get_dcf()
    ref = orc.pipe_menon(tn, N, 4.0, lut, iters=10)
    e = rel_l2(got, ref)
    print(f"[pipe-menon dim={dim}] 10 iterations, FP32 GPU vs FP64 oracle rel-L2 = {e:.2e}", flush=True)
    assert e <= 1e-6          # ten FP32 iterations of a ratio update (measured ~1.5e-7)


@pytest.mark.parametrize("dim", [2, 3])
def test_dcf_identity_scale(ctx, orc, dim):
    """fr_dcf_identity_scale: alpha = sum(w) / |G w|^2 against the FP64
    oracle's gridding of w, the in-place scaling bitwise w * alpha in FP32,
    and a second call on the scaled DCF returns 1 (the identity holds)."""
    if dim == 2:
        tn, dcf, data, N, lut, s, p, g = setup(ctx, 2, 60, 128, 1, 64, 3)
    else:
        tn, dcf, data, N, lut, s, p, g = setup(ctx, 3, 300, 32, 1, 16, 3)
    hip.pipe_menon(p, s, iters=3, init_ones=True)
    w = (s.get_dcf() * np.float32(2.5)).astype(np.float32)
    s.set_dcf(w)
    a = hip.dcf_identity_scale(p, s)
    g64 = orc.grid_f64(tn, w, np.ones((1, tn.shape[0]), np.complex64), N, 4.0, lut, nthreads=8)[0]
    ref = float(np.sum(w, dtype=np.float64) / np.sum(np.abs(g64) ** 2))
    got = s.get_dcf()
    a2 = hip.dcf_identity_scale(p, s)
    print(f"[identity scale dim={dim}] alpha {a:.6f} (FP64 oracle {ref:.6f}), second call {a2:.7f}", flush=True)
    assert abs(a / ref - 1.0) <= 1e-5
    assert np.array_equal(got.view(np.int32), (w * np.float32(a)).astype(np.float32).view(np.int32))
    assert abs(a2 - 1.0) <= 1e-5


def test_dcf_identity_scale_sampled(ctx):
    """The identity scale from a sample of the grid (the light units of one
    run of four pairs in every k, the heavy units of every k-th rank; each
    stratum scaled by its units over its sampled units): within 0.5% of the
    whole-grid identity on a 3D kooshball (1-iteration and converged
    Pipe-Menon DCFs, stride 4 and 8); a plan without unit pairs takes the
    whole grid (alpha within FP32 summation-order noise)."""
    from fastrecon_amd import ldcf
    traj, _ = syn.kooshball3d(8000, 128)
    N = 128
    tn, _ = hip.normalize_traj(traj, N)
    s, p = ldcf.plan_for(ctx, tn, N, 1)
    s1, p1 = ldcf.plan_for(ctx, tn, N, 1, dict(tile_writes=2))
    for iters in (1, 10):
        hip.pipe_menon(p, s, iters, True)
        w = s.get_dcf()
        a = {}
        for k in (1, 4, 8):
            s.set_dcf(w)
            a[k] = hip.dcf_identity_scale(p, s, k)
        s1.set_dcf(w)
        a_nopairs = hip.dcf_identity_scale(p1, s1, 8)
        print(f"[identity sampled, PM-{iters}] alpha whole {a[1]:.6f}, stride 4 {a[4]:.6f}, stride 8 {a[8]:.6f}, "
              f"no pairs {a_nopairs:.6f}", flush=True)
        assert abs(a[4] / a[1] - 1) <= 5e-3 and abs(a[8] / a[1] - 1) <= 5e-3
        assert a[4] != a[1] and a[8] != a[1]          # the sampled path ran
        assert abs(a_nopairs / a[1] - 1) <= 1e-5
    for h in (p1, s1, p, s):
        h.close()


def test_identity_scale_keeps_the_plan_current(ctx):
    """After the identity scale (whole grid and sampled) on a position-weight
    plan, which scales the DCF and the plan's copy of it together, the next
    gridding pass equals a fresh plan's pass on the scaled DCF, bit for bit;
    stride 0 is refused."""
    tn, dcf, data, N, lut, s, p, g = setup(ctx, 3, 2000, 64, 3, 32, 8)
    hip.pipe_menon(p, s, iters=2, init_ones=True)
    hip.grid_execute(p, s, g)                # the plan's DCF copy is now current
    for k in (1, 8):
        hip.dcf_identity_scale(p, s, k)
        hip.grid_execute(p, s, g)
        got = g.download_all()
        s2 = hip.Samples(ctx, 3, tn.shape[0], 3)
        s2.set_traj(tn)
        s2.set_dcf(s.get_dcf())
        s2.set_coils(data)
        p2 = hip.Plan(ctx, s2, N, 4.0, lut)
        g2 = hip.Grids(ctx, 3, N, 3)
        hip.grid_execute(p2, s2, g2)
        want = g2.download_all()
        print(f"[identity keeps plan current, stride {k}] equal {np.array_equal(got.view(np.int64), want.view(np.int64))}",
              flush=True)
        assert np.array_equal(got.view(np.int64), want.view(np.int64))
        for h in (g2, p2, s2):
            h.close()
    with pytest.raises(hip.HipError):
        hip.dcf_identity_scale(p, s, 0)
    for h in (g, p, s):
        h.close()


@pytest.mark.parametrize("dim,N,C", [(2, 64, 4), (3, 32, 3)])
def test_pils_ops_bitwise(ctx, orc, dim, N, C):
    """lowFilter, normalize, crop and SOS(map) are element-wise FP32 with the
    reference's operation order: bit-identical to the oracle."""
    rng = np.random.default_rng(17)
    cells = N ** dim
    d = (rng.standard_normal((C, cells)) + 1j * rng.standard_normal((C, cells))).astype(np.complex64)
    m = (rng.standard_normal((C, cells)) + 1j * rng.standard_normal((C, cells))).astype(np.complex64)
    g = hip.Grids(ctx, dim, N, C)
    for c in range(C):
        g.upload(c, d[c])
    hip.lowfilter(ctx, g, 7)
    assert np.array_equal(g.download_all(), orc.lowfilter(dim, N, d, 7))
    for c in range(C):
        g.upload(c, d[c])
    hip.normalize(ctx, g)
    assert np.array_equal(g.download_all(), orc.normalize(dim, N, d))

    rc = (N // 2, N // 2 + 2, N // 2 if dim == 3 else 1)
    for c in range(C):
        g.upload(c, d[c])
    out = hip.crop(ctx, g, rc).to_host(np.complex64).reshape(C, -1)
    full = d.reshape((C,) + (N,) * dim)
    x0, y0 = (N - rc[0]) // 2, (N - rc[1]) // 2
    if dim == 3:
        z0 = (N - rc[2]) // 2
        ref = full[:, z0:z0 + rc[2], y0:y0 + rc[1], x0:x0 + rc[0]]
    else:
        ref = full[:, y0:y0 + rc[1], x0:x0 + rc[0]]
    assert np.array_equal(out, ref.reshape(C, -1))

    gm = hip.Grids(ctx, dim, N, C)
    for c in range(C):
        gm.upload(c, m[c])
    got = hip.sos_map(ctx, g, gm, rc, shift=False).to_host(np.complex64)
    assert np.array_equal(got, orc.sos_map(dim, N, d, m, rc))
    # fftshift folded into the reads == explicit shift of both inputs
    got_s = hip.sos_map(ctx, g, gm, rc, shift=True).to_host(np.complex64)
    ref_s = orc.sos_map(dim, N, orc.fftshift(dim, N, d), orc.fftshift(dim, N, m), rc)
    assert np.array_equal(got_s, ref_s)
    # grids_copy is a device copy of every coil
    g2 = hip.Grids(ctx, dim, N, C)
    hip.grids_copy(ctx, g, g2)
    assert np.array_equal(g2.download_all(), d)


def _rand_grids(ctx, dim, N, C, seed, frames=0):
    rng = np.random.default_rng(seed)
    g = hip.Grids(ctx, dim, N, C, frames=frames)
    cells = N ** dim * max(frames, 1)
    data = (rng.standard_normal((C, cells), np.float32)
            + 1j * rng.standard_normal((C, cells), np.float32)).astype(np.complex64)
    for c in range(C):
        g.upload(c, data[c])
    return g, data


@pytest.mark.parametrize("dim,N,C,rc,shift,deapod,frames", [
    (3, 256, 3, (128, 128, 128), True, False, 0),
    (3, 256, 2, (96, 112, 100), True, True, 0),
    (3, 256, 2, (128, 128, 128), False, False, 0),
    (2, 512, 3, (256, 256, 1), True, False, 0),
    (2, 512, 2, (192, 160, 1), False, True, 0),
    (2, 256, 3, (128, 128, 1), True, False, 5),
    (3, 384, 2, (256, 256, 256), True, True, 0),    # N = 384: radices 4, 4, 4, 2, 3
    (3, 384, 1, (192, 160, 176), False, False, 0),
    (2, 384, 3, (192, 192, 1), True, False, 4),
    (2, 384, 2, (256, 240, 1), False, True, 0),
    (3, 128, 2, (64, 64, 64), True, False, 0),      # not a cropped-path size: the FFT + SOS fallback
])
def test_fft_sos_matches_fft_then_sos(ctx, dim, N, C, rc, shift, deapod, frames):
    """fr_fft_sos (cropped transforms for N = 256, 384, 512) equals fr_fft_execute
    followed by fr_sos on the same grids to FP32 rounding."""
    g, data = _rand_grids(ctx, dim, N, C, 31 + N + C, frames)
    batch = C * max(frames, 1)
    f = hip.FFT(ctx, dim, N, batch, +1)
    rcz = rc[2] if dim == 3 else max(frames, 1)
    dp = None
    if deapod:
        rng = np.random.default_rng(5)
        dmap = rng.uniform(0.5, 2.0, rc[0] * rc[1] * (rc[2] if dim == 3 else 1)).astype(np.float32)
        dp = hip.DeviceBuffer(ctx, 4 * dmap.size)
        dp.from_host(dmap)
    got = hip.fft_sos(f, g, rc, shift, dp).to_host(np.float32)
    for c in range(C):
        g.upload(c, data[c])
    f(g)
    ref = hip.sos(ctx, g, rc, shift, dp).to_host(np.float32)
    assert got.size == rc[0] * rc[1] * rcz
    e = rel_l2(got, ref)
    print(f"[fft_sos {dim}D N={N} C={C} rc={rc} shift={shift} deapod={deapod} frames={frames}] "
          f"rel-L2 vs FFT+SOS {e:.2e}", flush=True)
    assert e <= 1e-6
    for h in (f, g):
        h.close()


def test_fft_sos_against_oracle(ctx, orc):
    """The cropped path against the FP64 oracle FFT, fftshift and SOS."""
    N, C, rc = 256, 1, (128, 128, 128)
    g, data = _rand_grids(ctx, 3, N, C, 77)
    f = hip.FFT(ctx, 3, N, C, +1)
    got = hip.fft_sos(f, g, rc, True).to_host(np.float32)
    img = orc.fftshift(3, N, orc.fft_c2c(3, N, data, +1))
    ref = orc.sos(3, N, img, rc)
    e = rel_l2(got, ref)
    print(f"[fft_sos vs FP64 oracle] rel-L2 {e:.2e}", flush=True)
    assert e <= TOL


def test_fft_sos_384_against_oracle(ctx, orc):
    """The mixed-radix (4, 4, 4, 2, 3) N = 384 transform against the FP64
    oracle FFT, fftshift and SOS, 2D with three coils."""
    N, C, rc = 384, 3, (256, 256, 1)
    g, data = _rand_grids(ctx, 2, N, C, 78)
    f = hip.FFT(ctx, 2, N, C, +1)
    got = hip.fft_sos(f, g, rc, True).to_host(np.float32)
    img = orc.fftshift(2, N, orc.fft_c2c(2, N, data, +1))
    ref = orc.sos(2, N, img, rc)
    e = rel_l2(got, ref)
    print(f"[fft_sos N=384 vs FP64 oracle] rel-L2 {e:.2e}", flush=True)
    assert e <= TOL
