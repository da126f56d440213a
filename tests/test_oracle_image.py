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
