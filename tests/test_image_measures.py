"""The image measures of the learned-DCF evaluation (fastrecon_amd/ldcf.py:
nrmse, nrmse_scaled, psf_sidelobe) against closed forms.

- nrmse(a, b) = |a - b| / |b|: zero for a = b, |c - 1| for a = c b, and
  |e| / |b| for a = b + e with e orthogonal to b.
- nrmse_scaled(a, b) = min_c |c a - b| / |b| = |sin(angle(a, b))|: zero for
  any positive multiple of b, sin(theta) for a built at angle theta to b.
- psf_sidelobe(p, 2) = max |p| outside the 2-pixel ball around the peak,
  over the peak: for the separable sinc(x/2) sinc(y/2) sinc(z/2) sampled on
  integers (zeros on every even offset) the largest value outside the ball
  sits at (3, 0, 0): |sinc(3/2)| = 2 / (3 pi); points at distance exactly 2
  are inside the ball.
"""
import numpy as np
import pytest

from fastrecon_amd import ldcf


def rand(n, seed):
    return np.random.default_rng(seed).standard_normal(n)


def orthogonal_pair(n, seed):
    """b and a unit vector e orthogonal to b."""
    b = rand(n, seed)
    e = rand(n, seed + 1)
    e -= (e @ b) / (b @ b) * b
    return b, e / np.linalg.norm(e)


def test_nrmse_closed_forms():
    b, e = orthogonal_pair(4096, 1)
    nb = np.linalg.norm(b)
    assert ldcf.nrmse(b, b) == 0.0
    for c in (0.0, 0.5, 1.02, 3.0):
        assert ldcf.nrmse(c * b, b) == pytest.approx(abs(c - 1.0), rel=1e-12, abs=1e-15)
    for r in (1e-3, 0.01, 0.4):
        assert ldcf.nrmse(b + r * nb * e, b) == pytest.approx(r, rel=1e-10)
    # not symmetric: the second argument is the reference
    assert ldcf.nrmse(2 * b, b) == pytest.approx(1.0) and ldcf.nrmse(b, 2 * b) == pytest.approx(0.5)


@pytest.mark.parametrize("theta", [0.0, 1e-3, 0.01, 0.3, 1.2])
def test_nrmse_scaled_is_the_sine_of_the_angle(theta):
    b, e = orthogonal_pair(4096, 7)
    bu = b / np.linalg.norm(b)
    a = 5.0 * (np.cos(theta) * bu + np.sin(theta) * e)      # any length, angle theta to b
    assert ldcf.nrmse_scaled(a, b) == pytest.approx(np.sin(theta), rel=1e-9, abs=1e-12)
    # invariant to the scale of the first argument
    assert ldcf.nrmse_scaled(0.01 * a, b) == pytest.approx(ldcf.nrmse_scaled(a, b), rel=1e-9, abs=1e-12)
    # never above the unscaled measure
    assert ldcf.nrmse_scaled(a, b) <= ldcf.nrmse(a, b) + 1e-12


def sinc_psf(n, dim):
    x = np.arange(n) - n // 2
    s = np.sinc(x / 2.0)
    if dim == 2:
        return np.outer(s, s)
    return s[:, None, None] * s[None, :, None] * s[None, None, :]


@pytest.mark.parametrize("dim", [2, 3])
def test_psf_sidelobe_of_a_sampled_sinc(dim):
    p = sinc_psf(33, dim)
    assert ldcf.psf_sidelobe(p) == pytest.approx(2.0 / (3.0 * np.pi), rel=1e-12)
    # scale and sign of the PSF do not matter
    assert ldcf.psf_sidelobe(-4.0 * p) == pytest.approx(2.0 / (3.0 * np.pi), rel=1e-12)


def test_psf_sidelobe_ball_boundary_and_off_centre_peak():
    p = np.zeros((16, 16, 16))
    c = (5, 9, 4)                     # the peak need not be at the centre
    p[c] = 2.0
    p[c[0] + 2, c[1], c[2]] = 1.5     # distance exactly 2: inside the ball, ignored
    p[c[0] + 1, c[1] + 1, c[2] + 1] = 1.9   # distance sqrt(3): inside
    p[c[0], c[1] - 2, c[2] + 1] = 0.3       # distance sqrt(5): the sidelobe
    p[0, 0, 15] = -0.1
    assert ldcf.psf_sidelobe(p) == pytest.approx(0.15, rel=1e-12)
    assert ldcf.psf_sidelobe(p, radius=3) == pytest.approx(0.05, rel=1e-12)
