"""Learned splitting: train the subtree-size MLP on the solver's own node
counts, then use it to split and order the work queue of large solves.

Training data: every prefix (all row-0 columns) of N=13..17 at depths 4..6
(the generator drops a prefix whose next row has no free cell). The pools
come from the device generator; each label is log2(1 + nodes below the
prefix) as K1 counts it (per-prefix output). Everything runs through the C
ABI; nothing here counts on the CPU.

The queue order and the automatic split threshold (8x the level's mean
predicted size) use sizes relative to the level, so prediction_error reports
R^2 both as is and with each side's mean removed. With these 12 features
(row, free columns, attacked cells, free cells of the next 8 rows) both
stay near 0.65 at N=18 depth 6: a gradient-boosted regressor fitted to the
same features and labels reaches about the same (0.74 on held-out training
boards, 0.63 at N=18), so the features, not the network, set that limit.
"""
import time

import numpy as np

from . import DeviceBuffer, MIRROR_WEIGHT, make_tune

TRAIN_SPECS = [(n, d) for n in (13, 14, 15, 16, 17) for d in (4, 5, 6)]


def build_dataset(ctx, specs=TRAIN_SPECS):
    """Device features (count x 16 halves) and labels (float32) of every
    prefix in the full pools of `specs`, plus the total sample count."""
    feats, labels = [], []
    for n, d in specs:
        buf, cnt = ctx.pool_generate_device(n, d, 0)
        try:
            _, _, pn = ctx.count_pool_device(n, d, buf, cnt, per_prefix=True)
            fb = ctx.features_pool(n, d, buf, cnt)
            try:
                feats.append(fb.to_host(np.float16, cnt * 16).reshape(cnt, 16))
            finally:
                fb.free()
        finally:
            buf.free()
        labels.append(np.log2(pn.astype(np.float64) + 1.0).astype(np.float32))
    X = np.concatenate(feats)
    Y = np.concatenate(labels)
    return X, Y


def train(ctx, X, Y, epochs=12, batch=4096, lr=2e-3, seed=20261203, val_frac=0.2, model_seed=1):
    """Train a fresh model on (X, Y); returns (model, info)."""
    m = ctx.mlp(seed=model_seed)
    dx = DeviceBuffer.from_host(ctx, X)
    dy = DeviceBuffer.from_host(ctx, Y)
    try:
        _, v0 = m.train(dx, dy, len(Y), epochs=0, seed=seed, val_frac=val_frac)
        t0 = time.perf_counter()
        tr, v1 = m.train(dx, dy, len(Y), epochs=epochs, batch=batch, lr=lr, seed=seed,
                         val_frac=val_frac)
        dt = time.perf_counter() - t0
    finally:
        dx.free()
        dy.free()
    info = {"samples": int(len(Y)), "train_mse": tr, "val_mse": v1, "val_mse_untrained": v0,
            "label_var": float(np.var(Y)), "val_r2": 1.0 - v1 / float(np.var(Y)),
            "epochs": epochs, "batch": batch, "train_s": dt}
    return m, info


def prediction_error(ctx, model, n, depth):
    """Predicted vs K1-counted log2 subtree sizes of the half pool of n at
    `depth`: how the predictor is scored (no effect on any count)."""
    buf, cnt = ctx.pool_generate_device(n, depth, MIRROR_WEIGHT)
    try:
        _, _, pn = ctx.count_pool_device(n, depth, buf, cnt, per_prefix=True)
        fb = ctx.features_pool(n, depth, buf, cnt)
        try:
            pred = model.predict(fb, cnt)
        finally:
            fb.free()
    finally:
        buf.free()
    y = np.log2(pn.astype(np.float64) + 1.0)
    p = pred.astype(np.float64)
    err = p - y
    r2 = 1.0 - float(np.mean(err ** 2)) / float(np.var(y))
    # relative to the level's mean on both sides (what ordering and the
    # automatic threshold use)
    rel = (p - p.mean()) - (y - y.mean())
    r2_rel = 1.0 - float(np.mean(rel ** 2)) / float(np.var(y))
    ry = np.argsort(np.argsort(y))
    rp = np.argsort(np.argsort(pred))
    spearman = float(np.corrcoef(ry, rp)[0, 1])
    # rank agreement of the largest subtrees (what LPT ordering needs)
    k = max(1, cnt // 100)
    top_true = set(np.argsort(-y)[:k].tolist())
    top_pred = set(np.argsort(-pred)[:k].tolist())
    return {"n": n, "depth": depth, "prefixes": int(cnt), "mae_log2": float(np.mean(np.abs(err))),
            "rmse_log2": float(np.sqrt(np.mean(err ** 2))), "r2": r2, "r2_relative": r2_rel,
            "spearman": spearman,
            "top1pct_overlap": len(top_true & top_pred) / k}


def tune_learned(model, **kw):
    return make_tune(model=model, **kw)
