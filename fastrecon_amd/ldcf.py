"""Learned density compensation (net-new: the reference has no DCF
estimation).

An 8-64-64-1 SiLU MLP maps per-sample trajectory features (radius, azimuth,
polar angle, neighbour counts within 1/2/4 cells, distance to the outermost
sample shell, same-cell count; computed on the device by fr_ldcf_features)
to log DCF.  Its targets are 10-iteration Pipe-Menon DCFs
that this package's own gridding and degridding kernels produce
(fr_dcf_pipe_menon), on a seeded set of variable-density 3D radial and cones
trajectories for a 128^3 image (256^3 grid).  Training runs on the device
with the package's own BF16 forward/backward kernels and Adam
(fr_mlp_train_step).  The iterative DCF stays the fallback and the reference:
image NRMSE against the iterative-DCF image and the PSF sidelobe level.
"""
import json
import os
import time

import numpy as np

from . import hip
from . import synthetic as syn

KW, OG, L = 4.0, 2.0, 512
IMAGE = 128
N_GRID = syn.grid_size(IMAGE, OG)
CHECKPOINT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ldcf_kooshball.npz")
# the 2D network: 256^2 images on the 512^2 grid (the 2D radial workload's
# geometry), trained on 2D radial and spiral trajectories
IMAGE_2D = 256
N_GRID_2D = syn.grid_size(IMAGE_2D, OG)
CHECKPOINT_2D = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ldcf_radial2d.npz")


# ------------------------------------------------------- quality measures
def nrmse(a, b):
    """||a - b|| / ||b|| (no scale fit)."""
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def nrmse_scaled(a, b):
    """min_c ||c a - b|| / ||b|| (optimal real scale)."""
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    c = float(a @ b) / float(a @ a)
    return float(np.linalg.norm(c * a - b) / np.linalg.norm(b))


def psf_sidelobe(psf, radius=2):
    """max |PSF| outside a ball of `radius` pixels around the peak, over the peak."""
    p = np.abs(np.asarray(psf, np.float64))
    idx = np.unravel_index(np.argmax(p), p.shape)
    grids = np.meshgrid(*[np.arange(n) - i for n, i in zip(p.shape, idx)], indexing="ij")
    r2 = sum(g * g for g in grids)
    return float(p[r2 > radius * radius].max() / p[idx])


# --------------------------------------------------------------- pipeline
def plan_for(ctx, traj_n, N, coils=1, tune=None):
    M, dim = traj_n.shape
    s = hip.Samples(ctx, dim, M, coils)
    s.set_traj(traj_n)
    s.set_dcf(np.ones(M, np.float32))
    lut, _ = hip.kb_lut(KW, OG, L)
    p = hip.Plan(ctx, s, N, KW, lut, tune)
    return s, p


def pipe_menon_dcf(ctx, traj_n, N, iters=10):
    s, p = plan_for(ctx, traj_n, N)
    hip.pipe_menon(p, s, iters, True)
    w = s.get_dcf()
    feat = hip.ldcf_features(ctx, s, N).to_host(np.float32).reshape(-1, 8)
    for h in (p, s):
        h.close()
    return w, feat


def recon_image(ctx, plan, samples, N, rc, fft=None, grids=None):
    """grid -> FFT (+1) -> fftshift/crop/SOS, device resident (the cropped
    transforms of fr_fft_sos); returns the host image."""
    dim = samples.dim
    g = grids or hip.Grids(ctx, dim, N, samples.C)
    f = fft or hip.FFT(ctx, dim, N, samples.C, +1)
    hip.grid_execute(plan, samples, g)
    img = hip.fft_sos(f, g, rc, shift=True)
    out = img.to_host(np.float32)
    img.close()
    if grids is None:
        g.close()
    if fft is None:
        f.close()
    return out


# ---------------------------------------------------------------- dataset
def stratified_rows(radius, count, seed, offset, bins=32):
    """`count` row indices drawn (seeded, with replacement) equally from
    `bins` radius bins: the few samples near the k-space centre, which carry
    most of an object's energy, are as well represented as the many outer
    ones."""
    r = np.asarray(radius, np.float64)
    edges = np.linspace(0.0, r.max() + 1e-9, bins + 1)
    b = np.clip(np.searchsorted(edges, r, side="right") - 1, 0, bins - 1)
    order = np.argsort(b, kind="stable")
    starts = np.searchsorted(b[order], np.arange(bins))
    ends = np.searchsorted(b[order], np.arange(bins), side="right")
    filled = [k for k in range(bins) if ends[k] > starts[k]]
    u = syn.uniform(np.arange(count, dtype=np.uint64) + np.uint64(offset), seed)
    which = np.array(filled)[np.arange(count) % len(filled)]
    lo, hi = starts[which], ends[which]
    return order[lo + np.minimum((u * (hi - lo)).astype(np.int64), hi - lo - 1)]


def training_rows(feat, w, per_traj, seed, n):
    """Row indices of trajectory n of the training set (features `feat`,
    Pipe-Menon DCF `w`), three seeded strata:
    - half stratified over the radius in cells (from f0 = log1p(r)/5), which
      follows the many outer samples that carry most of a noise-like image;
    - a quarter stratified over f0, which keeps the few central samples that
      carry most of an object's energy;
    - a quarter drawn with probability proportional to w^2, each sample's
      weight in the image error of white k-space data.
    On the 32-coil kooshball the w^2 stratum took the learned image from 8.0%
    to 5.4% NRMSE (phantom 7.4% to 6.2%) over radius strata alone, at equal
    training steps."""
    f0 = feat[:, 0]
    r = np.expm1(5.0 * f0.astype(np.float64))
    p2 = np.asarray(w, np.float64) ** 2
    p2 /= p2.sum()
    rng = np.random.default_rng(1000 + n)
    return np.concatenate([stratified_rows(r, per_traj // 2, seed + 1, per_traj * n),
                           stratified_rows(f0, per_traj // 4, seed + 2, per_traj * n),
                           rng.choice(len(w), per_traj // 4, p=p2)])


def usable_targets(w):
    """Samples whose Pipe-Menon weight is a usable training target: positive,
    finite and at most 100 times the trajectory's 99th percentile.  (A sample
    at the very edge of k-space whose kernel falls mostly outside the grid
    has an almost empty density, and its weight can exceed every other
    sample's by orders of magnitude; the w^2 stratum of training_rows would
    draw nothing else.)"""
    w = np.asarray(w, np.float64)
    good = np.isfinite(w) & (w > 0)
    if not good.any():
        return good
    return good & (w <= 100.0 * np.percentile(w[good], 99))


def build_dataset(ctx, count=200, per_traj=50000, scale=1.0, seed=7, iters=10, log=print, dim=3):
    """Features and Pipe-Menon targets of `count` seeded training trajectories
    (3D: ldcf_training_trajectories on the 256^3 grid; 2D:
    ldcf_training_trajectories_2d on the 512^2 grid), `per_traj` rows drawn
    (seeded) from each (training_rows)."""
    X, W = [], []
    t0 = time.time()
    gen = syn.ldcf_training_trajectories if dim == 3 else syn.ldcf_training_trajectories_2d
    N = N_GRID if dim == 3 else N_GRID_2D
    for n, (name, k) in enumerate(gen(count, seed, scale)):
        tn, _ = hip.normalize_traj(k, N)
        w, feat = pipe_menon_dcf(ctx, tn, N, iters)
        ok = np.flatnonzero(usable_targets(w))
        pick = ok[training_rows(feat[ok], w[ok], per_traj, seed, n)]
        X.append(feat[pick])
        W.append(w[pick])
        if log and (n % 20 == 0 or n == count - 1):
            log(f"[ldcf data] {n + 1}/{count} {name}: M={len(w)} w in [{w.min():.3g}, {w.max():.3g}], "
                f"{len(w) - len(ok)} unusable targets ({time.time() - t0:.1f} s)")
    X = np.concatenate(X).astype(np.float32)
    W = np.concatenate(W).astype(np.float64)
    return X, W


# ---------------------------------------------------------------- training
def _init_params(seed):
    """He-style initialisation of the flat parameter vector (FP32)."""
    rng = np.random.default_rng(seed)
    nf, h = 8, 64
    W1 = rng.standard_normal((nf, h)) * np.sqrt(2.0 / nf)
    W2 = rng.standard_normal((h, h)) * np.sqrt(2.0 / h)
    w3 = rng.standard_normal(h) * np.sqrt(1.0 / h)
    return np.concatenate([W1.ravel(), np.zeros(h), W2.ravel(), np.zeros(h), w3, [0.0]]).astype(np.float32)


def train(ctx, X, W, steps=3000, batch=65536, lr=3e-3, seed=11, log=print, log_every=500, base=-1):
    """Adam on the MSE of the normalised log DCF, all on the device; with a
    base feature k (3..5) the target is log(w (1 + n_k)) = log w + 8 f_k,
    the DCF relative to the inverse local density, and the network's head
    divides by 1 + n_k again (fr_mlp_set_base).  Returns (mlp, mu, sigma,
    losses)."""
    logw = np.log(np.maximum(W, 1e-30))
    if base >= 0:
        logw = logw + 8.0 * X[:, base].astype(np.float64)
    mu, sigma = float(logw.mean()), float(logw.std())
    t = ((logw - mu) / sigma).astype(np.float32)
    n = len(t)
    feat = hip.DeviceBuffer(ctx, X.nbytes)
    feat.from_host(np.ascontiguousarray(X, np.float32))
    targ = hip.DeviceBuffer(ctx, t.nbytes)
    targ.from_host(t)
    idx = hip.DeviceBuffer(ctx, 8 * n)
    mlp = hip.MLP(ctx)
    mlp.set_params(_init_params(seed), mu, sigma)
    mlp.base = base
    rng = np.random.default_rng(seed)
    losses = []
    per_epoch = max(1, n // batch)
    perm = None
    for step in range(steps):
        e, o = divmod(step, per_epoch)
        if o == 0:
            perm = rng.permutation(n).astype(np.int64)
            idx.from_host(perm)
        lr_t = lr * 0.5 * (1.0 + np.cos(np.pi * step / steps))          # cosine decay
        want = (step % log_every == 0) or step == steps - 1
        B = min(batch, n)
        l = mlp.train_step(feat, targ, B, lr_t, idx=hip.DevicePtr(idx, 8 * o * B), want_loss=want)
        if want:
            losses.append((step, l))
            if log:
                log(f"[ldcf train] step {step} loss {l:.5f} lr {lr_t:.2e}")
    for b in (feat, targ, idx):
        b.close()
    return mlp, mu, sigma, losses


def calibrate_scale(ctx, mlp, dim=3, count=20, seed=7, scale=1.0, log=None):
    """Least-squares global scale of the network's DCF on whole training
    trajectories: c = sum(w_hat w) / sum(w_hat^2) over every sample of the
    first `count` trajectories of the training set, w the 10-iteration
    Pipe-Menon DCF, w_hat the network's.  For white k-space data the image
    error of a DCF error dw is about sum(dw^2) over the samples, so c is the
    scale that minimises it on that set.  The head's mu becomes mu + log c
    (no trajectory outside the training set is used).  Returns c."""
    gen = syn.ldcf_training_trajectories if dim == 3 else syn.ldcf_training_trajectories_2d
    N = N_GRID if dim == 3 else N_GRID_2D
    num = den = 0.0
    for n, (name, k) in enumerate(gen(count, seed, scale)):
        tn, _ = hip.normalize_traj(k, N)
        s, p = plan_for(ctx, tn, N)
        hip.pipe_menon(p, s, 10, True)
        w = s.get_dcf().astype(np.float64)
        mlp.dcf_from_traj(s, N)
        wh = s.get_dcf().astype(np.float64)
        ok = usable_targets(w)
        num += float(wh[ok] @ w[ok])
        den += float(wh[ok] @ wh[ok])
        for h in (p, s):
            h.close()
        if log:
            log(f"[ldcf calibrate] {n + 1}/{count} {name}: running scale {num / den:.5f}")
    c = num / den
    params, mu, sigma = mlp.get_params()
    base = mlp.base
    mlp.set_params(params, mu + float(np.log(c)), sigma)
    mlp.base = base
    return c


def save(mlp, path=CHECKPOINT, meta=None):
    p, mu, sigma = mlp.get_params()
    np.savez(path, params=p, mu=np.float32(mu), sigma=np.float32(sigma), base=np.int32(mlp.base),
             meta=json.dumps(meta or {}))


def load(ctx, path=CHECKPOINT):
    z = np.load(path, allow_pickle=False)
    mlp = hip.MLP(ctx)
    mlp.set_params(z["params"], float(z["mu"]), float(z["sigma"]))
    mlp.base = int(z["base"]) if "base" in z.files else -1
    return mlp, json.loads(str(z["meta"]))


# the identity scale's sampling stride on large grids: |G w|^2 from one run of
# unit pairs in every 8
IDENTITY_STRIDE = 8


def identity_stride(dim, N):
    """IDENTITY_STRIDE on grids of 2^22 cells or more; 1 (the whole grid) on
    smaller ones, where the sampled estimate's extra kernels cost more than
    the single-channel pass they shorten (a 512^2 grid: 0.47 against 0.38 ms
    on MI355X)."""
    return IDENTITY_STRIDE if N ** dim >= (1 << 22) else 1


def learned_dcf(ctx, mlp, samples, N, scale_plan=None, tile_stride=None):
    """Features of the samples' (normalised) trajectory -> MLP -> samples' DCF
    (one fused pass, fr_ldcf_dcf); with `scale_plan`, then the global scale
    of the Pipe-Menon identity (fr_dcf_identity_scale on that plan, |G w|^2
    from a sample of the unit pairs with stride tile_stride, by default
    identity_stride(dim, N); 1: the whole grid).  Returns the scale applied
    (1.0 without)."""
    mlp.dcf_from_traj(samples, N)
    if scale_plan is None:
        return 1.0
    k = identity_stride(samples.dim, N) if tile_stride is None else tile_stride
    return hip.dcf_identity_scale(scale_plan, samples, k)


def evaluate(ctx, mlp, traj, data, N, rc, ramp=None, iters=10, reps=1, log=None, coils=None):
    """Learned vs iterative DCF on one acquisition (device resident after
    upload): end-to-end milliseconds of each recon from the plan (iterative:
    Pipe-Menon + grid + FFT + SOS, on the recon's plan and on a second plan
    shaped for single-channel gridding; learned: features + MLP + grid + FFT +
    SOS; learned + identity scale: one single-channel gridding pass more),
    the plan build (ratios with it included: a fresh trajectory pays it on
    both sides), image NRMSE of learned (and ramp) against iterative, PSF
    sidelobes.
    Times with torch's HIP events, so torch must have been imported before
    the library was loaded (the binding's module docstring says why)."""
    import torch  # device timing through HIP events on the context stream

    tn, _ = hip.normalize_traj(traj, N)
    M = tn.shape[0]
    C = coils if callable(data) else data.shape[0]
    t0 = time.perf_counter()
    s, p = plan_for(ctx, tn, N, C)
    ctx.sync()
    t_plan = (time.perf_counter() - t0) * 1e3
    if callable(data):
        data(s)                  # the caller fills the coil data (e.g. on the device)
    else:
        s.set_coils(data)
    g = hip.Grids(ctx, tn.shape[1], N, C)
    f = hip.FFT(ctx, tn.shape[1], N, C, +1)
    img = hip.DeviceBuffer(ctx, 4 * rc[0] * rc[1] * rc[2])
    stream = torch.cuda.ExternalStream(ctx.stream)

    def timed(fn):
        best = None
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            ctx.sync()
            t = e0.elapsed_time(e1)
            best = t if best is None else min(best, t)
        return best

    def recon():
        hip.grid_execute(p, s, g)
        hip.fft_sos(f, g, rc, True, None, img)

    def iterative():
        hip.pipe_menon(p, s, iters, True)
        recon()

    def learned():
        mlp.dcf_from_traj(s, N)
        recon()

    t_it = timed(iterative)
    im_it = img.to_host(np.float32)
    w_it = s.get_dcf()
    t_ml = timed(learned)
    im_ml = img.to_host(np.float32)
    w_ml = s.get_dcf()
    t_rc = timed(recon)

    def learned_refined():
        # the learned DCF as the start of one Pipe-Menon iteration
        mlp.dcf_from_traj(s, N)
        hip.pipe_menon(p, s, 1, False)
        recon()

    t_mr = timed(learned_refined)
    im_mr = img.to_host(np.float32)

    def learned_scaled():
        # the learned DCF with its global scale from the Pipe-Menon identity
        mlp.dcf_from_traj(s, N)
        hip.dcf_identity_scale(p, s)
        recon()

    t_ms = timed(learned_scaled)
    im_ms = img.to_host(np.float32)

    def learned_sampled():
        # the same with |G w|^2 estimated from a sample of the unit pairs (stride IDENTITY_STRIDE)
        mlp.dcf_from_traj(s, N)
        hip.dcf_identity_scale(p, s, IDENTITY_STRIDE)
        recon()

    t_mss = timed(learned_sampled)
    im_mss = img.to_host(np.float32)
    mlp.dcf_from_traj(s, N)
    alpha = hip.dcf_identity_scale(p, s)     # the scale the identity gives the learned DCF
    mlp.dcf_from_traj(s, N)
    alpha_s = hip.dcf_identity_scale(p, s, IDENTITY_STRIDE)

    # the iterative DCF on a second plan of the same samples with the unit
    # shape that single-channel gridding prefers (the default shape for one
    # coil), then the same recon: the fastest iterative path this library has
    lut, _ = hip.kb_lut(KW, OG, L)
    p1 = hip.Plan(ctx, s, N, KW, lut, dict(tile_x=4, tile_y=4, tile_z=2) if tn.shape[1] == 3
                  else dict(tile_x=8, tile_y=4))

    def iterative_own_plan():
        hip.pipe_menon(p1, s, iters, True)
        recon()

    t_io = timed(iterative_own_plan)
    im_io = img.to_host(np.float32)
    p1.close()
    # one Pipe-Menon iteration (1 / kernel density) for reference
    hip.pipe_menon(p, s, 1, True)
    recon()
    im_pm1 = img.to_host(np.float32)
    out = dict(ms_plan=t_plan, ms_iterative=t_it, ms_learned=t_ml, ms_recon_only=t_rc, speedup=t_it / t_ml,
               speedup_with_plan=(t_plan + t_it) / (t_plan + t_ml),
               ms_learned_scaled=t_ms, speedup_learned_scaled=t_it / t_ms,
               speedup_learned_scaled_with_plan=(t_plan + t_it) / (t_plan + t_ms),
               learned_identity_scale=alpha,
               nrmse_learned_identity_scaled=nrmse(im_ms, im_it),
               ms_learned_sampled_scale=t_mss, speedup_learned_sampled_scale=t_it / t_mss,
               speedup_learned_sampled_scale_with_plan=(t_plan + t_it) / (t_plan + t_mss),
               identity_stride=IDENTITY_STRIDE, learned_sampled_identity_scale=alpha_s,
               nrmse_learned_sampled_identity_scaled=nrmse(im_mss, im_it),
               nrmse_pm1=nrmse(im_pm1, im_it), nrmse_pm1_scaled=nrmse_scaled(im_pm1, im_it),
               nrmse_learned=nrmse(im_ml, im_it), nrmse_learned_scaled=nrmse_scaled(im_ml, im_it),
               ms_learned_pm1=t_mr, speedup_learned_pm1=t_it / t_mr,
               ms_iterative_own_plan=t_io, speedup_vs_iterative_own_plan=t_io / t_ml,
               nrmse_iterative_own_plan=nrmse(im_io, im_it),
               nrmse_learned_pm1=nrmse(im_mr, im_it), nrmse_learned_pm1_scaled=nrmse_scaled(im_mr, im_it),
               dcf_rel_rms=float(np.sqrt(np.mean((w_ml / np.maximum(w_it, 1e-30) - 1.0) ** 2))))
    if ramp is not None:
        s.set_dcf(ramp)
        recon()
        im_rp = img.to_host(np.float32)
        out.update(nrmse_ramp=nrmse(im_rp, im_it), nrmse_ramp_scaled=nrmse_scaled(im_rp, im_it))
        # the ramp with the same identity scale (no reference image needed)
        hip.dcf_identity_scale(p, s)
        recon()
        out.update(nrmse_ramp_identity_scaled=nrmse(img.to_host(np.float32), im_it))
    # PSF: all-ones data with each DCF, one channel
    ones = hip.Samples(ctx, tn.shape[1], M, 1)
    ones.set_traj(tn)
    ones.set_coils(np.ones((1, M), np.complex64))
    g1 = hip.Grids(ctx, tn.shape[1], N, 1)
    f1 = hip.FFT(ctx, tn.shape[1], N, 1, +1)
    for name, w in (("iterative", w_it), ("learned", w_ml)) + ((("ramp", ramp),) if ramp is not None else ()):
        ones.set_dcf(w)
        psf = recon_image(ctx, p, ones, N, rc, f1, g1)
        out[f"psf_sidelobe_{name}"] = psf_sidelobe(psf.reshape(rc[2], rc[1], rc[0]) if tn.shape[1] == 3
                                                   else psf.reshape(rc[1], rc[0]))
    for h in (g1, f1, ones, img, f, g, p, s):
        h.close()
    if log:
        log(out)
    return out


# Axis-aligned ellipsoids (value, semi-axes a b c, centre x y z) in units of
# the image half-width, after the 3D Shepp-Logan layout of Kak & Slaney.
PHANTOM = [
    (1.0, 0.69, 0.92, 0.81, 0.0, 0.0, 0.0),
    (-0.8, 0.6624, 0.874, 0.78, 0.0, -0.0184, 0.0),
    (-0.2, 0.11, 0.31, 0.22, 0.22, 0.0, 0.0),
    (-0.2, 0.16, 0.41, 0.28, -0.22, 0.0, 0.0),
    (0.1, 0.21, 0.25, 0.41, 0.0, 0.35, -0.15),
    (0.1, 0.046, 0.046, 0.05, 0.0, 0.1, 0.25),
    (0.1, 0.046, 0.046, 0.05, 0.0, -0.1, 0.25),
    (0.1, 0.046, 0.023, 0.05, -0.08, -0.605, 0.0),
    (0.1, 0.023, 0.023, 0.02, 0.0, -0.606, 0.0),
    (0.1, 0.023, 0.046, 0.02, 0.06, -0.605, 0.0),
]


def phantom_kspace(k, image):
    """Exact k-space of the ellipsoid phantom at points k (cycles per pixel,
    M x 3): the Fourier transform of an ellipsoid of semi-axes (a, b, c) is
    abc * 4 pi (sin q - q cos q) / q^3 with q = 2 pi |(a kx, b ky, c kz)|,
    times the phase of its centre."""
    k = np.asarray(k, np.float64)
    half = image / 2.0
    out = np.zeros(k.shape[0], np.complex128)
    for A, a, b, c, x0, y0, z0 in PHANTOM:
        a, b, c = a * half, b * half, c * half
        q = 2 * np.pi * np.sqrt((a * k[:, 0]) ** 2 + (b * k[:, 1]) ** 2 + (c * k[:, 2]) ** 2)
        with np.errstate(invalid="ignore", divide="ignore"):
            f = np.where(q > 1e-6, 4 * np.pi * (np.sin(q) - q * np.cos(q)) / np.maximum(q, 1e-30) ** 3,
                         4 * np.pi / 3)
        ph = np.exp(-2j * np.pi * (k[:, 0] * x0 * half + k[:, 1] * y0 * half + k[:, 2] * z0 * half))
        out += A * a * b * c * f * ph
    return out.astype(np.complex64)


def evaluate_phantom(ctx, mlp, traj, N, rc, image, ramp=None, iters=10):
    """Single-coil phantom reconstructions with the iterative, learned and
    ramp DCFs: NRMSE of learned and ramp against iterative."""
    tn, _ = hip.normalize_traj(traj, N)
    s, p = plan_for(ctx, tn, N, 1)
    s.set_coils(phantom_kspace(traj, image)[None, :])
    f = hip.FFT(ctx, 3, N, 1, +1)
    g = hip.Grids(ctx, 3, N, 1)
    hip.pipe_menon(p, s, iters, True)
    im_it = recon_image(ctx, p, s, N, rc, f, g)
    learned_dcf(ctx, mlp, s, N)
    im_ml = recon_image(ctx, p, s, N, rc, f, g)
    hip.dcf_identity_scale(p, s)
    im_mi = recon_image(ctx, p, s, N, rc, f, g)
    learned_dcf(ctx, mlp, s, N, scale_plan=p)
    im_ms = recon_image(ctx, p, s, N, rc, f, g)
    out = dict(nrmse_learned=nrmse(im_ml, im_it), nrmse_learned_scaled=nrmse_scaled(im_ml, im_it),
               nrmse_learned_identity_scaled=nrmse(im_mi, im_it),
               nrmse_learned_sampled_identity_scaled=nrmse(im_ms, im_it))
    if ramp is not None:
        s.set_dcf(ramp)
        im_rp = recon_image(ctx, p, s, N, rc, f, g)
        out.update(nrmse_ramp=nrmse(im_rp, im_it), nrmse_ramp_scaled=nrmse_scaled(im_rp, im_it))
        hip.dcf_identity_scale(p, s)
        out.update(nrmse_ramp_identity_scaled=nrmse(recon_image(ctx, p, s, N, rc, f, g), im_it))
    for h in (g, f, p, s):
        h.close()
    return out
