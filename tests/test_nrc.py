"""Radiance cache: the MFMA kernels against the float64 numpy restatement
(oracle/mlp_ref.py), the restatement against finite differences, training
behaviour, and what the trained network gains inside the renderer."""
import numpy as np
import pytest

import mlp_ref
from conftest import relmse

BMIN, BMAX = (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)


def random_queries(n, seed=0):
    rng = np.random.default_rng(seed)
    f = np.zeros((n, 16), np.float32)
    f[:, 0:3] = rng.uniform(-1, 1, (n, 3))
    d = rng.normal(size=(n, 3))
    f[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    f[:, 6:9] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    f[:, 9] = rng.uniform(0, 1, n)
    f[:, 10:13] = rng.uniform(0.1, 0.9, (n, 3))
    return f


def random_params(seed=3, scale=1.0):
    rng = np.random.default_rng(seed)
    W = rng.uniform(-1, 1, (5, 64, 64)) * np.sqrt(6 / 64) * scale
    W[4][:, 3:] = 0
    b = rng.uniform(-0.1, 0.1, (5, 64))
    b[4][3:] = 0
    return mlp_ref.pack(W, b).astype(np.float32)


def test_reference_backward_matches_finite_differences():
    """Pins the restatement itself: analytic gradient == central differences (float64)."""
    rng = np.random.default_rng(1)
    p = random_params().astype(np.float64)
    X = mlp_ref.encode(random_queries(8), BMIN, BMAX)
    t = rng.uniform(0, 2, (8, 3))
    loss, g = mlp_ref.loss_and_grad(p, X, t, fp16=False)
    H, y = mlp_ref.forward(p, X, fp16=False)
    den = y * y + 0.01  # the loss holds the denominator constant (stop-gradient)
    f = lambda q: np.sum((mlp_ref.forward(q, X, fp16=False)[1] - t) ** 2 / den) / (3 * len(X))
    idx = np.concatenate([rng.choice(5 * 64 * 64, 40, replace=False), 5 * 64 * 64 + rng.choice(5 * 64, 10)])
    for i in idx:
        if i < 5 * 64 * 64 and (i // 4096) == 4 and (i % 64) >= 3:
            continue
        if i >= 5 * 64 * 64 and (i - 5 * 64 * 64) // 64 == 4 and (i - 5 * 64 * 64) % 64 >= 3:
            continue
        e = 1e-6
        q1, q2 = p.copy(), p.copy()
        q1[i] += e
        q2[i] -= e
        fd = (f(q1) - f(q2)) / (2 * e)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


@pytest.mark.gpu
def test_mfma_forward_matches_reference(gpu_available):
    import crystal2_amd as c2
    cache = c2.RadianceCache(BMIN, BMAX, batch_capacity=4096)
    p = random_params()
    cache.set_params(p)
    q = random_queries(1000)
    y = cache.infer(q)
    _, yr = mlp_ref.forward(p, mlp_ref.encode(q, BMIN, BMAX), fp16=True)
    yr = np.maximum(yr, 0)
    err = np.abs(y - yr).max() / (np.abs(yr).max() + 1e-6)
    assert err < 2e-2, err


@pytest.mark.gpu
def test_mfma_gradients_match_reference(gpu_available):
    import crystal2_amd as c2
    cache = c2.RadianceCache(BMIN, BMAX, batch_capacity=4096)
    p = random_params(scale=0.7)
    cache.set_params(p)
    q = random_queries(3000, seed=5)
    t = np.random.default_rng(6).uniform(0, 1, (3000, 3)).astype(np.float32)
    g, loss = cache.train_step(q, t, apply=False)
    lr, gr = mlp_ref.loss_and_grad(p, mlp_ref.encode(q, BMIN, BMAX), t, fp16=True)
    assert abs(loss - lr) <= 2e-2 * abs(lr)
    for L in range(5):
        a = g[L * 4096:(L + 1) * 4096].reshape(64, 64)
        b = gr[L * 4096:(L + 1) * 4096].reshape(64, 64)
        if L == 4:
            a, b = a[:, :3], b[:, :3]
        rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-12)
        assert rel < 3e-2, (L, rel)
    gb, gbr = g[5 * 4096:].reshape(5, 64), gr[5 * 4096:].reshape(5, 64)
    gb[4][3:] = gbr[4][3:] = 0
    assert np.linalg.norm(gb - gbr) / np.linalg.norm(gbr) < 3e-2


@pytest.mark.gpu
def test_training_fits_a_smooth_field(gpu_available):
    import crystal2_amd as c2
    cache = c2.RadianceCache(BMIN, BMAX, batch_capacity=1 << 14, learning_rate=1e-2)
    q = random_queries(1 << 14, seed=9)
    t = np.stack([0.5 + 0.4 * np.sin(3 * q[:, 0]), 0.5 + 0.4 * np.cos(2 * q[:, 1]), 0.3 + 0.2 * q[:, 2]], 1)
    t = t.astype(np.float32)
    losses = [cache.train_step(q, t)[1] for _ in range(300)]
    assert losses[-1] < 0.1 * losses[0], (losses[0], losses[-1])
    test = random_queries(2000, seed=10)
    tt = np.stack([0.5 + 0.4 * np.sin(3 * test[:, 0]), 0.5 + 0.4 * np.cos(2 * test[:, 1]),
                   0.3 + 0.2 * test[:, 2]], 1)
    assert np.mean(np.abs(cache.infer(test) - tt)) < 0.05


@pytest.mark.gpu
def test_cache_gain_in_renderer(gpu_available):
    """Paths cut at bounce 2: the trained cache must beat the same cut without it
    (a zero network = plain truncation) against a high-spp plain ground truth."""
    import bench
    import crystal2_amd as c2
    sc = bench.build_scene(n_spheres=60, level=2, seed=4)
    W, H, depth = 160, 90, 8
    truth = c2.WavefrontPathIntegrator(sc, W, H, 256, depth, film="f64")
    truth.render()
    ref = truth.film()
    lo, hi = bench.scene_bounds(sc)
    results = {}
    for name in ("trained", "zero"):
        if name == "trained":
            cache = c2.RadianceCache(lo, hi, seed=2)
        else:  # the same truncation without a network: zero weights, no learning
            cache = c2.RadianceCache(lo, hi, seed=2, learning_rate=0.0)
            cache.set_params(np.zeros(c2.RadianceCache.PARAMS, np.float32))
        pt = c2.WavefrontPathIntegrator(sc, W, H, 64, depth, film="kahan")
        pt.set_cache(cache, cache_bounce=2, train_ratio=16)
        pt.render(0, 64)
        pt.sync()
        f = pt.film()
        assert np.all(np.isfinite(f))
        results[name] = relmse(f, ref)
    assert results["trained"] < 0.7 * results["zero"], results
