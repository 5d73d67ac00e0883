"""The autotuner sweeps expansion depth x waves per CU x LDS stack rows; every
trial must give identical counts, and the winner is a member of the grid."""
import pytest

import nqm
from nqm import autotune, workloads

pytestmark = pytest.mark.gpu


def test_sweep_small_batch(ctx, tmp_path):
    solves = workloads.mixed_batch(count=64, n_lo=8, n_hi=14)
    best, rows = autotune.sweep(ctx, solves, depths=(3, 5), waves=(4, 16), stack_rows=(5, 7),
                                repeats=1)
    assert len(rows) == 8
    assert (best.depth, best.waves_per_cu, best.stack_rows) in {(d, w, r) for d, w, r, _ in rows}
    assert all(ms > 0 for *_, ms in rows)
    path = str(tmp_path / "cache.json")
    autotune.store(ctx, "small64", best, rows, path=path)
    t = autotune.cached_tune(ctx, "small64", path=path)
    assert (t.depth, t.waves_per_cu, t.stack_rows) == (best.depth, best.waves_per_cu, best.stack_rows)
    c1, _, _ = ctx.solve_batch(solves, tune=t)
    c0, _, _ = ctx.solve_batch(solves)
    assert (c1 == c0).all()
