"""Learned density compensation on the GPU through the C ABI: features
against the oracle's brute-force definition, the BF16 MLP forward and its
gradient against the FP64 oracle (same BF16 rounding points), determinism,
training on the device, and what the learned DCF gains over the analytic
ramp DCF in a reconstruction measured against the Pipe-Menon (iterative) DCF.

Bars: neighbour counts exact; angles 1e-5; forward |dy| <= 1e-2 (a BF16
rounding flip of one hidden unit moves y by ~1e-3) with median <= 1e-5;
gradient relative L2 <= 1e-3."""
import numpy as np
import pytest

import ldcf_oracle as lo
from fastrecon_amd import hip, ldcf
from fastrecon_amd import synthetic as syn

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = hip.Context(0)
    yield c
    c.close()


def samples_for(ctx, tn):
    M, dim = tn.shape
    s = hip.Samples(ctx, dim, M, 1)
    s.set_traj(tn)
    s.set_dcf(np.ones(M, np.float32))
    return s


@pytest.mark.parametrize("dim,N,M,near", [(3, 32, 3000, 0.5), (2, 48, 2500, 0.5), (3, 5, 400, 0.5),
                                          (2, 6, 300, 0.5), (2, 6000, 2000, 0.5), (3, 330, 1500, 0.5),
                                          (3, 40, 12000, 0.95)])
def test_features_match_oracle(ctx, dim, N, M, near):
    """A fraction `near` of the points gathered around the grid centre, the
    rest uniform.  Grids of at most 8192 tiles (16^3 cells in 3D, 64^2 or
    128^2 in 2D) count cells per tile in LDS; N = 5 and 6 are smaller than
    one tile; N = 6000 in 2D uses 128^2 tiles and has more radius shells
    (8487) than the count pass keeps in LDS, so it counts them in memory;
    N = 330 in 3D has 9261 tiles and counts cells with one atomic per sample
    (a central box in LDS); N = 40 with 12000 points puts more than 8192 of
    them in the central tile, which is counted in two chunks."""
    rng = np.random.default_rng(dim + 10)
    Mn = int(M * near)
    t = np.concatenate([rng.normal(N / 2, 2.5, (Mn, dim)), rng.uniform(-0.7, N - 0.3, (M - Mn, dim))])
    t = t.astype(np.float32)
    s = samples_for(ctx, t)
    f = hip.ldcf_features(ctx, s, N).to_host(np.float32).reshape(M, 8)
    ref = lo.features(t, N)
    print(f"[features dim={dim}] max |df| = {np.abs(f - ref).max(axis=0)}", flush=True)
    for q, r in enumerate((1, 2, 4)):
        n_gpu = np.rint(np.expm1(f[:, 3 + q].astype(np.float64) * 8)).astype(np.int64)
        assert np.array_equal(n_gpu, lo.neighbour_counts(t, N, r))
    np.testing.assert_allclose(f[:, 0], ref[:, 0], rtol=2e-6, atol=1e-6)
    np.testing.assert_allclose(f[:, 1:3], ref[:, 1:3], atol=1e-5)
    # outer-shell distance: exact FP32 radii and maximum
    np.testing.assert_allclose(f[:, 6], ref[:, 6], rtol=0, atol=2e-6)
    assert np.sum(f[:, 6] == 0.0) >= 1
    # shell density: exact FP32 radii and counts, an FP32 shell volume
    np.testing.assert_allclose(f[:, 7], ref[:, 7], rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("dim,N,M", [(3, 32, 3000), (3, 40, 37), (2, 48, 2561), (3, 64, 64 * 41), (3, 6, 500),
                                     (3, 330, 700)])
def test_fused_dcf_equals_features_then_network(ctx, dim, N, M):
    """fr_ldcf_dcf (features in LDS, no feature rows) is bitwise the DCF of
    fr_ldcf_features + fr_mlp_dcf; M below, at and off multiples of 64, and a
    second call on a smaller grid reuses the kept workspace (N = 330 counts
    cells per sample, its second grid of 168 per tile)."""
    rng = np.random.default_rng(dim * 100 + M)
    t = np.concatenate([rng.normal(N / 2, 2.5, (M // 2, dim)), rng.uniform(-0.7, N - 0.3, (M - M // 2, dim))])
    t = t.astype(np.float32)
    net = hip.MLP(ctx)
    net.set_params(random_net(dim + M), 0.3, 0.7)
    for n in (N, N // 2 + 3):
        s = samples_for(ctx, t)
        feat = hip.ldcf_features(ctx, s, n)
        net.dcf(feat, s)
        want = s.get_dcf()
        s.set_dcf(np.zeros(M, np.float32))
        net.dcf_from_traj(s, n)
        got = s.get_dcf()
        print(f"[fused dcf dim={dim} N={n} M={M}] equal {np.array_equal(got.view(np.int32), want.view(np.int32))}",
              flush=True)
        np.testing.assert_array_equal(got.view(np.int32), want.view(np.int32))
        feat.close()
        s.close()
    net.close()


def random_net(seed):
    rng = np.random.default_rng(seed)
    return (lo.init_params(seed) + 0.05 * rng.standard_normal(lo.NPARAM)).astype(np.float32)


def upload(ctx, a):
    a = np.ascontiguousarray(a)
    b = hip.DeviceBuffer(ctx, max(a.nbytes, 4))
    b.from_host(a)
    return b


def test_forward_matches_oracle(ctx):
    rng = np.random.default_rng(3)
    M = 5000 + 7                     # not a multiple of 16
    X = np.zeros((M, 8), np.float32)
    X[:, :] = rng.uniform(-1.5, 1.5, (M, 8))
    p = random_net(3)
    net = hip.MLP(ctx)
    net.set_params(p, 0.0, 1.0)
    y = net.forward(upload(ctx, X), M).to_host(np.float32)
    ref, _ = lo.forward(p, X, emulate=True)
    d = np.abs(y - ref)
    print(f"[mlp forward] max |dy| {d.max():.2e}, median {np.median(d):.2e}", flush=True)
    assert np.median(d) <= 1e-5 and d.max() <= 1e-2
    # the DCF head is exp(mu + sigma y), divided by 1 + n_k with a base feature
    net.set_params(p, 0.25, 0.5)
    s = hip.Samples(ctx, 3, M, 1)
    net.dcf(upload(ctx, X), s)
    w = s.get_dcf()
    np.testing.assert_allclose(w, np.exp(0.25 + 0.5 * y), rtol=2e-6)
    for k in (3, 4, 5):
        net.base = k
        assert net.base == k
        net.dcf(upload(ctx, X), s)
        np.testing.assert_allclose(s.get_dcf(), lo.dcf_head(y, X, 0.25, 0.5, k), rtol=4e-6)
    with pytest.raises(hip.HipError):
        net.base = 7
    net.base = -1


def test_gradient_matches_oracle_and_is_deterministic(ctx):
    rng = np.random.default_rng(4)
    B = 4096 + 5
    X = np.zeros((B, 8), np.float32)
    X[:, :] = rng.uniform(-1.5, 1.5, (B, 8))
    t = rng.standard_normal(B).astype(np.float32)
    p = random_net(4)
    net = hip.MLP(ctx)
    net.set_params(p, 0.0, 1.0)
    fx, ft = upload(ctx, X), upload(ctx, t)
    g1, l1 = net.grad(fx, ft, B)
    g2, l2 = net.grad(fx, ft, B)
    assert np.array_equal(g1, g2) and l1 == l2
    Lr, gr = lo.loss_grad(p, X, t, emulate=True)
    e = np.linalg.norm(g1 - gr) / np.linalg.norm(gr)
    print(f"[mlp grad] loss {l1:.6f} vs {Lr:.6f}, grad rel-L2 {e:.2e}", flush=True)
    assert abs(l1 - Lr) <= 1e-4 * abs(Lr) and e <= 1e-3
    # a row subset through an index array equals the gathered batch
    idx = rng.permutation(B)[:1000].astype(np.int64)
    gi, li = net.grad(fx, ft, 1000, idx=upload(ctx, idx))
    gs, ls = net.grad(upload(ctx, X[idx]), upload(ctx, t[idx]), 1000)
    assert np.array_equal(gi, gs) and li == ls


def test_training_reduces_loss(ctx):
    """A smooth target of the features is learned on the device."""
    rng = np.random.default_rng(5)
    n = 65536
    X = np.zeros((n, 8), np.float32)
    X[:, :] = rng.uniform(0, 1, (n, 8))
    W = np.exp(2.0 * X[:, 0] ** 2 - X[:, 3] + 0.5 * np.sin(3 * X[:, 1]))
    net, mu, sigma, losses = ldcf.train(ctx, X, W, steps=400, batch=8192, lr=3e-3, log=None, log_every=100)
    print(f"[ldcf train] losses {losses}", flush=True)
    assert losses[-1][1] < 0.05 * losses[0][1]


def small_eval(ctx, mlp, spokes=3000, readout=128, coils=4, image=48):
    """Images of a kooshball with the iterative, learned and ramp DCFs."""
    N = syn.grid_size(image, 2.0)
    traj, ramp = syn.kooshball3d(spokes, readout)
    tn, _ = hip.normalize_traj(traj, N)
    M = tn.shape[0]
    data = syn.coil_data(coils, M, 9)
    s, p = ldcf.plan_for(ctx, tn, N, coils)
    s.set_coils(data)
    rc = (image, image, image)
    hip.pipe_menon(p, s, 10, True)
    img_it = ldcf.recon_image(ctx, p, s, N, rc)
    ldcf.learned_dcf(ctx, mlp, s, N)
    img_ml = ldcf.recon_image(ctx, p, s, N, rc)
    s.set_dcf(ramp)
    img_rp = ldcf.recon_image(ctx, p, s, N, rc)
    return img_it, img_ml, img_rp


def test_learned_dcf_gain_over_ramp(ctx):
    """Train a small network on a small seeded set of training trajectories
    (same generator as the shipped one, scaled down) and check it beats the
    analytic ramp DCF against the iterative-DCF image."""
    X, W = ldcf.build_dataset(ctx, count=12, per_traj=20000, scale=0.05, log=None)
    net, mu, sigma, losses = ldcf.train(ctx, X, W, steps=600, batch=16384, lr=3e-3, log=None)
    img_it, img_ml, img_rp = small_eval(ctx, net)
    e_ml, e_rp = ldcf.nrmse_scaled(img_ml, img_it), ldcf.nrmse_scaled(img_rp, img_it)
    print(f"[ldcf gain] loss {losses[0][1]:.4f} -> {losses[-1][1]:.4f}; image NRMSE vs iterative (scaled): "
          f"learned {e_ml:.4f}, ramp {e_rp:.4f}", flush=True)
    assert e_ml < e_rp
