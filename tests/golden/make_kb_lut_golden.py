"""Writes kb_lut_golden.npz: the reference KB LUT (ConvKernel.cpp:7-30)
evaluated with scipy.special.i0 (Cephes Chebyshev expansion -- an I0
implementation independent of the oracle's power series) under numpy
float32/float64 semantics matching the reference's promotions.

Run from the repo root:  python tests/golden/make_kb_lut_golden.py
"""
import os

import numpy as np
from scipy.special import i0

CASES = [(4.0, 2.0, 512), (4.0, 1.5, 512), (4.0, 2.0, 32), (3.0, 2.0, 512), (5.0, 1.25, 64)]


def lut(w, a, L):
    w = np.float32(w)
    a = np.float32(a)
    q = np.float32(np.float32(w * w) / np.float32(a * a))
    am = np.float64(a) - 0.5
    beta = np.float32(np.pi * np.sqrt(np.float64(q) * am * am - 0.8))
    dk = np.float32(np.float64(w) / 2.0 / np.float64(L - 1))
    out = np.zeros(L, np.float32)
    k0 = np.float32(1.0)
    for i in range(L):
        k = np.float32(dk * np.float32(i))
        t = np.float32(np.float32(np.float32(2.0) * k) / w)
        x = np.float64(beta) * np.sqrt(1.0 - np.float64(t) ** 2)
        v = np.float32(i0(x) / np.float64(w))
        if i == 0:
            k0 = v
        out[i] = np.float32(v / k0)
    return out, beta


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    arrs = {}
    for (w, a, L) in CASES:
        t, beta = lut(w, a, L)
        key = f"w{w:g}_og{a:g}_L{L}".replace(".", "p")
        arrs[key] = t
        arrs[key + "_beta"] = np.array([beta], np.float32)
    np.savez(os.path.join(here, "kb_lut_golden.npz"), **arrs)
    print("wrote", sorted(arrs))


if __name__ == "__main__":
    main()
