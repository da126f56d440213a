"""CPU oracle for the learned density compensation (net-new: the reference
has no DCF estimation).  TEST INFRASTRUCTURE ONLY: used by tests/ and never by
the product.

Definitions restated here, independently of the HIP kernels:

features(t, N) for normalised trajectory points t (grid units, as after
ReconData::normalizeTraj, ReconData.cpp:211-225), with cf = (N-1)/2, h = N/2:
  u_d   = (t_d - cf) / h                      (FP32)
  f0    = log1p(|u| N/2) / 5                  radius in cells, log-compressed
  f1    = atan2(u_y, u_x) / pi                azimuth
  f2    = acos(clamp(u_z/|u|)) / pi           polar (0.5 in 2D or at u = 0)
  cell  = clamp(floor(t_d + 0.5), 0, N-1)
  n_r   = #{samples j != i : max_d |cell_j - cell_i| <= r},  r = 1, 2, 4
  f3..5 = log1p(n_r) / 8
  rho   = |u| N/2 (FP32: sqrt, then times N/2), rho_max = max over samples
  f6    = min(rho_max - rho, 4) / 4           distance to the outermost shell
  b     = min(floor(2 rho), nb - 1), nb = ceil(N sqrt(dim)) + 2
  c_b   = #{samples j : b_j = b}              (this sample included)
  V_b   = (4 pi / 3)(hi^3 - lo^3) in 3D, pi (hi^2 - lo^2) in 2D, lo = b/2,
          hi = lo + 1/2                       cells of the radius shell
  f7    = log1p(c_b / V_b) / 8                samples per cell of the shell
The box counts are pinned by brute force O(M^2) here.

MLP 8 -> 64 -> 64 -> 1 with SiLU, BF16 weight matrices, FP32 biases:
  z1 = x W1 + b1, h1 = bf16(silu(z1)), z2 = h1 W2 + b2, h2 = silu(z2),
  y  = h2 w3 + b3,
with W1, W2, w3 the BF16-rounded values.  DCF head (dcf_head): w = exp(mu +
sigma y), divided by 1 + n_k = exp(8 f_k) for a network with base feature k.
The backward restates the chain rule (straight through the BF16 rounding of
h1) and is pinned by central finite differences of the same function in FP64
without rounding.
"""
import numpy as np

NF = 8
H = 64


def bf16_round(a):
    """FP32 -> BF16 (round to nearest even) -> FP32 values."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def features(tn, N):
    t = np.asarray(tn, np.float32)
    M, d = t.shape
    cf = np.float32(0.5) * np.float32(N - 1)
    h = np.float32(0.5) * np.float32(N)
    u = (t - cf) / h
    if d == 3:
        r = np.sqrt((u[:, 0] * u[:, 0] + u[:, 1] * u[:, 1]) + u[:, 2] * u[:, 2])
    else:
        r = np.sqrt(u[:, 0] * u[:, 0] + u[:, 1] * u[:, 1])
    f = np.zeros((M, NF), np.float64)
    f[:, 0] = np.log1p(r.astype(np.float64) * float(h)) / 5.0
    f[:, 1] = np.arctan2(u[:, 1].astype(np.float64), u[:, 0].astype(np.float64)) / np.pi
    if d == 3:
        with np.errstate(invalid="ignore", divide="ignore"):
            c = np.clip(u[:, 2].astype(np.float64) / r.astype(np.float64), -1.0, 1.0)
        f[:, 2] = np.where(r > 0, np.arccos(c) / np.pi, 0.5)
    else:
        f[:, 2] = 0.5
    f[:, 3:6] = np.log1p(neighbour_counts(t, N, (1, 2, 4))) / 8.0
    rho = (r * h).astype(np.float32)
    f[:, 6] = np.minimum((rho.max() - rho).astype(np.float32), np.float32(4.0)).astype(np.float64) * 0.25
    f[:, 7] = shell_density(rho, N, d)
    return f


def shell_bins(N, dim):
    return int(np.ceil(N * np.sqrt(dim))) + 2


def shell_cells(b, dim):
    """Cells of the radius shell b/2 <= rho < b/2 + 1/2 (FP64)."""
    lo = np.asarray(b, np.float64) / 2.0
    hi = lo + 0.5
    return 4.0 * np.pi / 3.0 * (hi ** 3 - lo ** 3) if dim == 3 else np.pi * (hi ** 2 - lo ** 2)


def shell_density(rho, N, dim):
    """f7: log1p(samples per cell of each sample's radius shell) / 8."""
    nb = shell_bins(N, dim)
    b = np.clip(np.floor(2.0 * np.asarray(rho, np.float64)), 0, nb - 1).astype(np.int64)
    c = np.bincount(b, minlength=nb)
    return np.log1p(c[b] / shell_cells(b, dim)) / 8.0


def dcf_head(y, x, mu, sigma, base=-1):
    """The DCF from network outputs y of feature rows x."""
    e = mu + sigma * np.asarray(y, np.float64)
    if base >= 0:
        e = e - 8.0 * np.asarray(x, np.float64)[:, base]
    return np.exp(e)


def neighbour_counts(tn, N, r):
    """Other samples whose nearest cell lies within Chebyshev distance r of
    this sample's: every pair compared (blocks of 128 rows at a time).  r is
    one radius (an (M,) result) or a tuple of radii (an (M, len(r)) one)."""
    t = np.asarray(tn, np.float32)
    radii = np.atleast_1d(np.asarray(r, np.int64))
    cell = np.clip(np.floor(t + np.float32(0.5)), 0, N - 1).astype(np.int32)
    n = np.empty((len(cell), len(radii)), np.int64)
    for i in range(0, len(cell), 128):
        d = np.abs(cell[i:i + 128, None, 0] - cell[None, :, 0])
        for k in range(1, cell.shape[1]):
            np.maximum(d, np.abs(cell[i:i + 128, None, k] - cell[None, :, k]), out=d)
        for q, rr in enumerate(radii):
            n[i:i + 128, q] = np.count_nonzero(d <= rr, axis=1) - 1
    return n if np.ndim(r) else n[:, 0]


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def unpack(p):
    """Flat parameter vector (the C ABI's order) -> W1 (NF,64), b1, W2 (64,64), b2, w3 (64), b3."""
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + NF * H].reshape(NF, H); o += NF * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]; o += 1
    return W1, b1, W2, b2, w3, b3


NPARAM = NF * H + H + H * H + H + H + 1


def forward(p, x, emulate=True):
    """Returns y and the intermediates; emulate=True applies the kernel's
    BF16 rounding of the weight matrices and of h1."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    if emulate:
        W1, W2, w3 = (bf16_round(W).astype(np.float64) for W in (W1, W2, w3))
    x = np.asarray(x, np.float64)
    z1 = x @ W1 + b1
    h1 = silu(z1)
    if emulate:
        h1 = bf16_round(h1).astype(np.float64)
    z2 = h1 @ W2 + b2
    h2 = silu(z2)
    y = h2 @ w3 + b3
    return y, (x, z1, h1, z2, h2, W1, W2, w3)


def loss_grad(p, x, t, emulate=True):
    """Mean squared error L = mean((y - t)^2) and dL/dp (flat, C ABI order)."""
    y, (x, z1, h1, z2, h2, W1, W2, w3) = forward(p, x, emulate)
    B = len(y)
    r = y - np.asarray(t, np.float64)
    L = float(np.mean(r * r))
    dy = 2.0 * r / B
    dw3 = h2.T @ dy
    db3 = dy.sum()
    dz2 = (dy[:, None] * w3[None, :]) * dsilu(z2)
    dW2 = h1.T @ dz2
    db2 = dz2.sum(0)
    dh1 = dz2 @ W2.T
    dz1 = dh1 * dsilu(z1)
    dW1 = x.T @ dz1
    db1 = dz1.sum(0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dw3, [db3]])
    return L, g


def init_params(seed=0):
    """He-style initialisation, FP32 master values."""
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((NF, H)) * np.sqrt(2.0 / NF)
    W2 = rng.standard_normal((H, H)) * np.sqrt(2.0 / H)
    w3 = rng.standard_normal(H) * np.sqrt(1.0 / H)
    return np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), w3, [0.0]]).astype(np.float32)
