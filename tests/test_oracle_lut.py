"""Pins the oracle's KB LUT (ConvKernel.cpp:7-30) and I0."""
import math
import os

import numpy as np
import pytest
from scipy.special import i0 as sp_i0

GOLD = os.path.join(os.path.dirname(__file__), "golden", "kb_lut_golden.npz")


def _cases():
    g = np.load(GOLD)
    for k in g.files:
        if k.endswith("_beta"):
            continue
        w, og, L = k.split("_")
        yield k, float(w[1:].replace("p", ".")), float(og[2:].replace("p", ".")), int(L[1:])


def test_i0_against_scipy(orc):
    for x in np.linspace(0.0, 12.0, 241):
        assert orc.i0(x) == pytest.approx(sp_i0(x), rel=2e-15)


@pytest.mark.parametrize("case", list(_cases()), ids=lambda c: c[0])
def test_lut_matches_golden_bitwise(orc, case):
    key, w, og, L = case
    g = np.load(GOLD)
    lut, beta = orc.kb_lut(w, og, L)
    assert np.float32(beta) == g[key + "_beta"][0]
    np.testing.assert_array_equal(lut.view(np.int32), g[key].view(np.int32))


@pytest.mark.parametrize("og,beta_ref", [(2.0, 8.996152), (1.5, 7.8922853)])
def test_lut_closed_forms(orc, og, beta_ref):
    lut, beta = orc.kb_lut(4.0, og, 512)
    assert beta == pytest.approx(beta_ref, abs=5e-7)
    assert lut[0] == 1.0
    # at i = L-1, 2k/w == 1 exactly, so the entry is (1/w) / (I0(beta)/w)
    assert lut[-1] == pytest.approx(1.0 / sp_i0(beta), rel=1e-6)
    assert np.all(np.diff(lut) < 0)          # monotone decreasing kernel
    # every entry is the analytic KB profile I0(b sqrt(1-(2k/w)^2)) / I0(b)
    k = np.arange(512) * (4.0 / 2.0 / 511)
    ana = sp_i0(beta * np.sqrt(np.clip(1 - (2 * k / 4.0) ** 2, 0, None))) / sp_i0(beta)
    np.testing.assert_allclose(lut, ana, rtol=3e-6)


def test_lut_beta_formula_beatty(orc):
    # Beatty et al. 2005: beta = pi sqrt(w^2/a^2 (a - 1/2)^2 - 0.8)
    for w, a in [(4.0, 2.0), (4.0, 1.5), (3.0, 2.0), (6.0, 1.25)]:
        _, beta = orc.kb_lut(w, a, 32)
        assert beta == pytest.approx(math.pi * math.sqrt(w * w / (a * a) * (a - 0.5) ** 2 - 0.8), rel=1e-6)
