"""Argument errors through the C ABI: every case below is refused on the host,
before anything is launched, with a status code and a message naming the
problem (the C++ layer turns these into the reference's cerr + exit(1)).
Empty inputs are refused here; the reference's loops would produce an empty
or zero result from them."""
import numpy as np
import pytest

from fastrecon_amd import hip

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = hip.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("dim,M,coils,match", [
    (3, 0, 4, "sample count 0"),
    (4, 100, 4, "dim must be 2 or 3"),
    (3, 100, 0, "coil count 0"),
])
def test_samples_arguments(ctx, dim, M, coils, match):
    with pytest.raises(hip.HipError, match=match):
        hip.Samples(ctx, dim, M, coils)


def test_frame_without_samples(ctx):
    with pytest.raises(hip.HipError, match="frame 1 has no samples"):
        hip.Samples(ctx, 2, 150, 2, frame_samples=[100, 0, 50])


def test_fft_sign(ctx):
    with pytest.raises(hip.HipError, match="sign must be"):
        hip.FFT(ctx, 2, 64, 1, 2)


def test_fft_sos_arguments(ctx):
    g = hip.Grids(ctx, 2, 64, 2)
    f = hip.FFT(ctx, 2, 64, 2, +1)
    with pytest.raises(hip.HipError, match="crop 80x32x1 outside grid 64"):
        hip.fft_sos(f, g, (80, 32, 1), True)
    f3 = hip.FFT(ctx, 2, 64, 3, +1)
    with pytest.raises(hip.HipError, match="do not match the FFT plan"):
        hip.fft_sos(f3, g, (32, 32, 1), True)
    small = hip.DeviceBuffer(ctx, 4 * 16)
    with pytest.raises(hip.HipError, match="device allocation holds"):
        hip.fft_sos(f, g, (32, 32, 1), True, None, small)
    for h in (small, f3, f, g):
        h.close()


def test_odd_grid_shift(ctx):
    g = hip.Grids(ctx, 2, 63, 1)
    with pytest.raises(hip.HipError, match="even grid size"):
        hip.sos(ctx, g, (32, 32, 1), True)
    g.close()


def test_device_data_extent(ctx):
    """A coil-data buffer smaller than M * C complex64 values is refused
    before any copy (the extent of the device allocation is checked)."""
    M, C = 1000, 4
    s = hip.Samples(ctx, 3, M, C)
    buf = hip.DeviceBuffer(ctx, M * C * 4)          # half of what complex64 needs
    with pytest.raises(hip.HipError, match="device allocation holds"):
        s.set_data_device(buf.ptr, hip.FR_LAYOUT_SAMPLE_MAJOR)
    buf.close()
    s.close()
    # the context stays usable after refused calls
    s2 = hip.Samples(ctx, 3, M, C)
    s2.set_coils(np.ones((C, M), np.complex64))
    s2.close()
