"""The subtree-size MLP: features, forward and backward kernels against the
CPU FP32 reference (tests/mlp_ref.py), and training on solver-recorded data.

Tolerances: with the device's FP16 rounding points replicated on the CPU the
gradients agree to rtol 2e-3 (only summation order differs); against pure
FP32 math they agree to rtol 5e-2, the FP16 input rounding."""
import numpy as np
import pytest

import mlp_ref
import nqm
import oracle

pytestmark = pytest.mark.gpu


def _dataset(ctx, specs):
    feats, labels = [], []
    for n, d in specs:
        pool = oracle.pool_generate(n, d, mirror=False)
        buf = nqm.DeviceBuffer.from_host(ctx, pool)
        try:
            _, _, pn = ctx.count_pool_device(n, d, buf, pool.shape[0], per_prefix=True)
        finally:
            buf.free()
        feats.append(mlp_ref.features(n, pool))
        labels.append(np.log2(pn.astype(np.float64) + 1).astype(np.float32))
    return np.concatenate(feats), np.concatenate(labels)


def test_features_match_cpu(ctx):
    for n, d in ((10, 3), (14, 5), (17, 6)):
        pool = oracle.pool_generate(n, d, mirror=True)[:2000]
        buf = nqm.DeviceBuffer.from_host(ctx, pool)
        try:
            fb = ctx.features_pool(n, d, buf, pool.shape[0])
            got = fb.to_host(np.float16, pool.shape[0] * 16).reshape(-1, 16)
            fb.free()
        finally:
            buf.free()
        want = mlp_ref.features(n, pool).astype(np.float16)
        assert np.array_equal(got, want), (n, d)


def test_forward_and_gradient_match_cpu(ctx):
    rng = np.random.default_rng(3)
    m = ctx.mlp(seed=11)
    p = m.get_params()
    p[mlp_ref.IN * mlp_ref.HID:] += rng.normal(0, 0.05, p.size - mlp_ref.IN * mlp_ref.HID).astype(np.float32)
    m.set_params(p)
    x = np.zeros((1000, 16), np.float32)
    x[:, :12] = rng.uniform(0, 1.2, (1000, 12))
    x = x.astype(np.float16)
    t = rng.uniform(2, 20, 1000).astype(np.float32)
    dx = nqm.DeviceBuffer.from_host(m.ctx, x)
    dt = nqm.DeviceBuffer.from_host(m.ctx, t)
    try:
        y = m.predict(dx, 1000)
        g, loss = m.gradient(dx, dt, 1000)
    finally:
        dx.free()
        dt.free()
    y_ref, _ = mlp_ref.forward(p, x.astype(np.float32), match_device=True)
    assert np.allclose(y, y_ref, rtol=1e-3, atol=1e-3)
    g_ref, loss_ref = mlp_ref.gradient(p, x.astype(np.float32), t, match_device=True)
    assert abs(loss - loss_ref) <= 1e-3 * abs(loss_ref)
    scale = np.abs(g_ref).max()
    assert np.allclose(g, g_ref, rtol=2e-3, atol=2e-4 * scale)
    g32, _ = mlp_ref.gradient(p, x.astype(np.float32), t, match_device=False)
    assert np.allclose(g, g32, rtol=5e-2, atol=5e-3 * scale)
    # padded feature rows of W1 never move
    assert np.all(g[12 * 64:16 * 64] == 0)
    m.close()


def test_training_learns_subtree_sizes(ctx):
    X, Y = _dataset(ctx, [(11, 4), (12, 4), (12, 5), (13, 5)])
    m = ctx.mlp(seed=5)
    dx = nqm.DeviceBuffer.from_host(ctx, X.astype(np.float16))
    dy = nqm.DeviceBuffer.from_host(ctx, Y)
    try:
        _, v0 = m.train(dx, dy, len(Y), epochs=0)
        tr, v1 = m.train(dx, dy, len(Y), epochs=30, batch=1024, lr=3e-3)
    finally:
        dx.free()
        dy.free()
    var = float(np.var(Y))
    assert v1 < 0.25 * var, (v0, v1, var)   # explains > 75% of the label variance
    assert v1 < v0
    m.close()
