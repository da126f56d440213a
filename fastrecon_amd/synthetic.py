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


# ------------------------------------------- learned-DCF training trajectories
def random_rotation(seed):
    """A seeded rotation matrix (QR of a Gaussian matrix, sign-fixed)."""
    g = complex_gaussian(9, seed).real.reshape(3, 3)
    q, r = np.linalg.qr(g)
    q = q * np.sign(np.diag(r))[None, :]
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def radial3d_vd(spokes, readout, alpha=1.0, seed=0):
    """Full-diameter spokes on rotated Fibonacci directions with variable
    density along the spoke: k = 0.5 sign(s) |s|^alpha u, s = (j - R/2)/(R/2)."""
    u = fibonacci_dirs(spokes) @ random_rotation(seed).T
    s = (np.arange(readout, dtype=np.float64) - readout / 2) / (readout / 2)
    r = 0.5 * np.sign(s) * np.abs(s) ** alpha
    k = (r[None, :, None] * u[:, None, :]).reshape(-1, 3)
    return k.astype(np.float32)


def cones3d(cones, interleaves, readout, turns=4.0, alpha=1.0, seed=0):
    """Centre-out 3D cones: cone c has polar angle (c + 1/2) pi / cones and
    `interleaves` spirals k = 0.5 t^a (sin th cos ph, sin th sin ph, cos th),
    ph = 2 pi turns t^a + 2 pi i / interleaves + (seeded cone offset)."""
    t = np.arange(readout, dtype=np.float64) / readout
    ta = t ** alpha
    off = uniform(np.arange(cones, dtype=np.uint64), seed) * 2.0 * math.pi
    ks = []
    for c in range(cones):
        th = (c + 0.5) * math.pi / cones
        for i in range(interleaves):
            ph = 2.0 * math.pi * turns * ta + 2.0 * math.pi * i / interleaves + off[c]
            ks.append(np.stack([0.5 * ta * math.sin(th) * np.cos(ph), 0.5 * ta * math.sin(th) * np.sin(ph),
                                0.5 * ta * math.cos(th) * np.ones_like(ta)], axis=1))
    return np.concatenate(ks, 0).astype(np.float32)


def ldcf_training_trajectories(count=200, seed=7, scale=1.0):
    """The seeded training set: variable-density 3D radial and cones
    trajectories for a 128^3 image (256^3 grid).  `scale` shrinks the spoke /
    interleave counts (tests use small sets).  Yields (name, k[M,3])."""
    for n in range(count):
        r = uniform(np.arange(4, dtype=np.uint64) + 4 * n, seed)
        if n % 2 == 0:
            # half of the radial sets are uniform-density (alpha 1), the
            # common clinical kooshball; most read 256 samples per spoke
            spokes = max(16, int(scale * (4000 + r[0] * 56000)))
            readout = int(128 + 128 * (r[1] > 0.25))
            alpha = 1.0 if r[2] < 0.5 else 1.0 + 2.0 * (r[2] - 0.5)
            yield f"radial{n}", radial3d_vd(spokes, readout, alpha, seed=seed * 1000 + n)
        else:
            cones = max(4, int(scale * (16 + r[0] * 48)))
            inter = max(4, int(scale * (64 + r[1] * 448)))
            alpha = 1.0 + 1.0 * r[2]
            turns = 2.0 + 6.0 * r[3]
            yield f"cones{n}", cones3d(cones, inter, 256, turns, alpha, seed=seed * 1000 + n)


def radial2d_vd(spokes, readout, alpha=1.0, golden=True, seed=0):
    """Full-diameter 2D spokes, golden-angle or uniformly spaced angles with a
    seeded offset, variable density along the spoke: k = 0.5 sign(s) |s|^a
    (cos t, sin t), s = (j - R/2)/(R/2)."""
    n = np.arange(spokes, dtype=np.float64)
    off = float(uniform(np.arange(1, dtype=np.uint64), seed)[0]) * math.pi
    th = off + (n * GOLDEN_ANGLE if golden else n * math.pi / spokes)
    s = (np.arange(readout, dtype=np.float64) - readout / 2) / (readout / 2)
    r = 0.5 * np.sign(s) * np.abs(s) ** alpha
    k = np.stack([(r[None, :] * np.cos(th)[:, None]).reshape(-1), (r[None, :] * np.sin(th)[:, None]).reshape(-1)], 1)
    return k.astype(np.float32)


def ldcf_training_trajectories_2d(count=200, seed=7, scale=1.0):
    """The seeded 2D training set for a 256^2 image (512^2 grid): half
    radial (golden-angle or uniform angles, 100-800 spokes of 256 or 512
    samples, half of them uniform density), half variable-density spirals
    (8-64 interleaves of 1024-4096 samples).  Yields (name, k[M,2])."""
    for n in range(count):
        r = uniform(np.arange(4, dtype=np.uint64) + 4 * n, seed + 1)
        if n % 2 == 0:
            spokes = max(16, int(scale * (100 + r[0] * 700)))
            readout = int(256 + 256 * (r[1] > 0.3))
            alpha = 1.0 if r[2] < 0.5 else 1.0 + 2.0 * (r[2] - 0.5)
            yield f"radial2d_{n}", radial2d_vd(spokes, readout, alpha, golden=r[3] < 0.5, seed=seed * 1000 + n)
        else:
            inter = max(4, int(scale * (8 + r[0] * 56)))
            samples = int(1024 * 2 ** int(r[1] * 3))
            alpha = 1.0 + 1.5 * r[2]
            turns = 4.0 + 20.0 * r[3]
            k, _ = spiral2d(inter, samples, alpha, turns)
            yield f"spiral2d_{n}", k


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


def spiral_frames(frames=64, samples=4096, seed=4, lo=8, hi=48):
    """The batched uneven 2D spiral series: frame f has a seeded number of
    interleaves in [lo, hi].  Returns (list of (k, dcf) per frame, counts)."""
    counts = spiral_frame_counts(frames, lo, hi, seed)
    return [spiral2d(int(n), samples) for n in counts], counts
