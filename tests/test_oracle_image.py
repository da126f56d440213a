"""Pins the oracle's FFT (FFT.cpp:44-47,83: in place, sign +1, unnormalised),
fftshift (ImageFilter.cpp:231-280) and SOS (ImageFilter.cpp:159-209)."""
import numpy as np
import pytest


@pytest.mark.parametrize("dim,N", [(2, 12), (2, 30), (3, 8), (3, 12), (2, 7)])
def test_fft_matches_numpy_backward_unnormalised(orc, dim, N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal((2, N ** dim)) + 1j * rng.standard_normal((2, N ** dim))
    got = orc.fft_c2c(dim, N, x, +1)
    shp = (2,) + (N,) * dim
    ref = np.fft.ifftn(x.reshape(shp), axes=tuple(range(1, dim + 1))) * N ** dim
    np.testing.assert_allclose(got.reshape(shp), ref, rtol=1e-12, atol=1e-10)
    fwd = orc.fft_c2c(dim, N, x, -1)
    np.testing.assert_allclose(fwd.reshape(shp), np.fft.fftn(x.reshape(shp), axes=tuple(range(1, dim + 1))),
                               rtol=1e-12, atol=1e-10)


def test_fft_delta_sign_and_parseval(orc):
    N = 16
    x = np.zeros((1, N * N), np.complex128)
    x[0, 1] = 1.0                                  # delta at x=1
    X = orc.fft_c2c(2, N, x, +1).reshape(N, N)
    kx = np.arange(N)
    np.testing.assert_allclose(X[0], np.exp(+2j * np.pi * kx / N), atol=1e-13)   # sign +1
    rng = np.random.default_rng(0)
    y = rng.standard_normal((1, N * N)) + 0j
    Y = orc.fft_c2c(2, N, y, +1)
    assert (np.abs(Y) ** 2).sum() == pytest.approx(N * N * (np.abs(y) ** 2).sum(), rel=1e-12)


@pytest.mark.parametrize("dim,N", [(2, 8), (2, 16), (3, 8), (3, 6)])
def test_fftshift_matches_numpy_and_is_involution(orc, dim, N):
    rng = np.random.default_rng(1)
    x = (rng.standard_normal((3, N ** dim)) + 1j * rng.standard_normal((3, N ** dim))).astype(np.complex64)
    s = orc.fftshift(dim, N, x)
    shp = (3,) + (N,) * dim
    np.testing.assert_array_equal(s.reshape(shp), np.fft.fftshift(x.reshape(shp), axes=tuple(range(1, dim + 1))))
    np.testing.assert_array_equal(orc.fftshift(dim, N, s), x)


@pytest.mark.parametrize("dim,N,rc", [(2, 16, (8, 8, 1)), (3, 12, (6, 6, 6)), (3, 10, (4, 6, 8))])
def test_sos_closed_form(orc, dim, N, rc):
    rng = np.random.default_rng(2)
    C = 3
    x = (rng.standard_normal((C, N ** dim)) + 1j * rng.standard_normal((C, N ** dim))).astype(np.complex64)
    got = orc.sos(dim, N, x, rc)
    v = x.reshape((C,) + (N,) * dim)
    if dim == 2:
        x0, y0 = (N - rc[0]) // 2, (N - rc[1]) // 2
        crop = v[:, y0:y0 + rc[1], x0:x0 + rc[0]]
    else:
        x0, y0, z0 = (N - rc[0]) // 2, (N - rc[1]) // 2, (N - rc[2]) // 2
        crop = v[:, z0:z0 + rc[2], y0:y0 + rc[1], x0:x0 + rc[0]]
    ref = np.sqrt((np.abs(crop.astype(np.complex128)) ** 2).sum(0)).reshape(-1)
    np.testing.assert_allclose(got, ref, rtol=2e-7)
    got64 = orc.sos(dim, N, x.astype(np.complex128), rc)
    np.testing.assert_allclose(got64, ref, rtol=1e-14)


@pytest.mark.parametrize("dim,N,res", [(2, 64, 22), (3, 32, 22), (2, 48, 5)])
def test_lowfilter_window(orc, dim, N, res):
    """lowFilter is data * exp(-floor(r/att*100)/100) with r the integer
    squared distance to size/2, att = res^2/2, and zero past idx 2000
    (res 5: att 12.5 is where the float quotient moves idx across integers)."""
    rng = np.random.default_rng(3)
    C = 2
    cells = N ** dim
    d = (rng.standard_normal((C, cells)) + 1j * rng.standard_normal((C, cells))).astype(np.complex64)
    out = orc.lowfilter(dim, N, d, res)
    ax = np.arange(N) - N // 2
    if dim == 3:
        z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
        r = (x * x + y * y + z * z).ravel()
    else:
        y, x = np.meshgrid(ax, ax, indexing="ij")
        r = (x * x + y * y).ravel()
    # the quotient is a float (int / float in the reference), the scale double
    q = r.astype(np.float32) / np.float32(res * res / 2.0)
    idx = np.floor(q.astype(np.float64) * 100.0).astype(np.int64)
    w = np.where(idx >= 2000, 0.0, np.exp(-idx / 100.0))
    ref = d.astype(np.complex128) * w
    # expf of a float argument: |arg| <= 20 carries up to 20 * 2^-24 relative
    np.testing.assert_allclose(out, ref, rtol=2e-6, atol=0)
    centre = np.flatnonzero(r == 0)
    assert np.array_equal(out[:, centre], d[:, centre])
    assert np.all(out[:, idx >= 2000] == 0)


@pytest.mark.parametrize("dim,N,C", [(2, 32, 4), (3, 16, 3)])
def test_normalize_unit_coil_norm(orc, dim, N, C):
    rng = np.random.default_rng(4)
    cells = N ** dim
    d = (rng.standard_normal((C, cells)) + 1j * rng.standard_normal((C, cells))).astype(np.complex64)
    out = orc.normalize(dim, N, d)
    norm = np.sum(np.abs(out.astype(np.complex128)) ** 2, axis=0)
    np.testing.assert_allclose(norm, 1.0, rtol=1e-6)
    # the per-voxel scale is real and common to all coils
    ratio = out.astype(np.complex128) / d.astype(np.complex128)
    np.testing.assert_allclose(ratio.imag, 0, atol=1e-6)
    np.testing.assert_allclose(ratio, np.broadcast_to(ratio[:1], ratio.shape), rtol=1e-6)
    np.testing.assert_allclose(ratio[0].real, 1 / np.sqrt(np.sum(np.abs(d.astype(np.complex128)) ** 2, 0)),
                               rtol=1e-6)


@pytest.mark.parametrize("dim,N,rc", [(2, 32, (16, 16, 1)), (3, 16, (8, 8, 8))])
def test_sos_map_closed_forms(orc, dim, N, rc):
    """SOS(map) with map == data is the SOS squared (real); scaling the map
    by a complex a scales the result by conj(a); the crop is centred."""
    rng = np.random.default_rng(5)
    C = 3
    cells = N ** dim
    d = (rng.standard_normal((C, cells)) + 1j * rng.standard_normal((C, cells))).astype(np.complex64)
    m = (rng.standard_normal((C, cells)) + 1j * rng.standard_normal((C, cells))).astype(np.complex64)
    same = orc.sos_map(dim, N, d, d, rc)
    s = orc.sos(dim, N, d, rc)
    np.testing.assert_allclose(same.real, s.astype(np.float64) ** 2, rtol=1e-6)
    assert np.all(same.imag == 0)
    a = np.complex64(0.5 - 2.0j)
    r1 = orc.sos_map(dim, N, d, m, rc).astype(np.complex128)
    r2 = orc.sos_map(dim, N, d, (m * a).astype(np.complex64), rc).astype(np.complex128)
    np.testing.assert_allclose(r2, np.conj(a) * r1, rtol=1e-5, atol=1e-5)
    # brute force in complex128 over the centred crop
    shape = (N,) * dim
    dd = d.reshape((C,) + shape).astype(np.complex128)
    mm = m.reshape((C,) + shape).astype(np.complex128)
    rcx, rcy, rcz = rc
    x0, y0 = (N - rcx) // 2, (N - rcy) // 2
    if dim == 3:
        z0 = (N - rcz) // 2
        sl = (slice(None), slice(z0, z0 + rcz), slice(y0, y0 + rcy), slice(x0, x0 + rcx))
    else:
        sl = (slice(None), slice(y0, y0 + rcy), slice(x0, x0 + rcx))
    ref = np.sum(dd[sl] * np.conj(mm[sl]), axis=0).ravel()
    np.testing.assert_allclose(r1, ref, rtol=1e-5, atol=1e-5)
