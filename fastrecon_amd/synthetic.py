"""Seeded synthetic trajectories and k-space data for the benchmark configs.

Every generator is deterministic given its seed and identical on every host:
random numbers come from a counter-based SplitMix64 stream, uniforms are the
top 53 bits, normals come from Box-Muller in FP64 and are cast to float32.

Trajectories are returned in the reference's *input* units (k in
[-0.5, 0.5)), exactly as a user would load them from files; the gridding
plan normalises them with the reference's component-0 bounds map.

Workloads (grid N = image * oversampling):
  radial2d     2D golden-angle radial, 402 spokes x 512, 8 coils, 512^2 grid
  kooshball3d  3D radial on Fibonacci-sphere directions, 51,200 spokes x 256,
               32 coils, 256^3 grid
  kooshball3d_fp16   100,000 spokes x 384, 16 coils, 384^3 grid, FP16 samples
  spiral2d_frames    64 frames of 8..48 variable-density spiral interleaves
               x 4096, 12 coils, 384^2 grid
"""
import math

import numpy as np

GOLDEN_ANGLE = math.pi * (math.sqrt(5.0) - 1.0) / 2.0          # 111.2461 deg
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(index, seed):
    """SplitMix64 output for counter values `index` (uint64 array)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.asarray(index, np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(index, seed):
    """Uniform doubles in [0, 1) from the top 53 bits."""
    return (splitmix64(index, seed) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def complex_gaussian(n, seed, offset=0):
    """n complex samples with re, im ~ N(0, 1) (sigma 1 per component)."""
    i = np.arange(offset, offset + n, dtype=np.uint64)
    u1 = uniform(2 * i, seed)
    u2 = uniform(2 * i + np.uint64(1), seed)
    r = np.sqrt(-2.0 * np.log1p(-u1))
    th = 2.0 * math.pi * u2
    out = np.empty(n, np.complex64)
    out.real = (r * np.cos(th)).astype(np.float32)
    out.imag = (r * np.sin(th)).astype(np.float32)
    return out


def coil_data(C, M, seed):
    """(C, M) complex64; coil c uses stream offset c*M."""
    return np.stack([complex_gaussian(M, seed, c * M) for c in range(C)])


# ---------------------------------------------------------------- 2D radial
def radial2d(spokes=402, readout=512):
    """Golden-angle radial: k = ((j - R/2)/R) * (cos t_n, sin t_n)."""
    n = np.arange(spokes, dtype=np.float64)
    th = n * GOLDEN_ANGLE
    j = (np.arange(readout, dtype=np.float64) - readout / 2) / readout
    kx = (j[None, :] * np.cos(th)[:, None]).reshape(-1)
    ky = (j[None, :] * np.sin(th)[:, None]).reshape(-1)
    traj = np.stack([kx, ky], axis=1).astype(np.float32)
    dcf = np.sqrt(kx * kx + ky * ky).astype(np.float32)          # ramp |k|
    return traj, dcf


# ------------------------------------------------------------ 3D kooshball
def fibonacci_dirs(n):
    i = np.arange(n, dtype=np.float64)
    z = 1.0 - (2.0 * i + 1.0) / n
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def kooshball3d(spokes=51200, readout=256):
    """Full-diameter spokes on Fibonacci-sphere directions; DCF = |k|^2."""
    u = fibonacci_dirs(spokes)
    j = (np.arange(readout, dtype=np.float64) - readout / 2) / readout
    k = (j[None, :, None] * u[:, None, :]).reshape(-1, 3)
    traj = k.astype(np.float32)
    dcf = (k * k).sum(1).astype(np.float32)
    return traj, dcf


# ------------------------------------------------------- 2D spiral frames
SPIRAL_ALPHA = 1.5
SPIRAL_TURNS = 16.0


def spiral2d(interleaves, samples=4096, alpha=SPIRAL_ALPHA, turns=SPIRAL_TURNS):
    """Variable-density Archimedean spiral k(t) = 0.5 t^a e^{i(2 pi T t^a + 2 pi l / n)}."""
    t = np.arange(samples, dtype=np.float64) / samples
    ta = t ** alpha
    ks = []
    for l in range(interleaves):
        ph = 2.0 * math.pi * turns * ta + 2.0 * math.pi * l / interleaves
        ks.append(np.stack([0.5 * ta * np.cos(ph), 0.5 * ta * np.sin(ph)], axis=1))
    k = np.concatenate(ks, 0)
    traj = k.astype(np.float32)
    dcf = np.sqrt((k * k).sum(1)).astype(np.float32)
    return traj, dcf


def spiral_frame_counts(frames=64, lo=8, hi=48, seed=4):
    u = uniform(np.arange(frames, dtype=np.uint64), seed)
    return (lo + np.floor(u * (hi - lo + 1))).astype(np.int64)


# ------------------------------------------------------------ config table
CONFIGS = {
    "radial2d": dict(dim=2, image=256, og=2.0, kwidth=4.0, L=512, coils=8, seed=1),
    "kooshball3d": dict(dim=3, image=128, og=2.0, kwidth=4.0, L=512, coils=32, seed=2),
    "kooshball3d_fp16": dict(dim=3, image=256, og=1.5, kwidth=4.0, L=512, coils=16, seed=3),
    "spiral2d_frames": dict(dim=2, image=192, og=2.0, kwidth=4.0, L=512, coils=12, seed=4),
}


def grid_size(image, og):
    """GridLut::plan: N = unsigned(reconSize * og) (GridLut.cpp:45)."""
    return int(np.float32(image) * np.float32(og))
