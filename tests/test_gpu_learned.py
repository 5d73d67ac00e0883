"""Learned splitting keeps every count exact: the model only moves work
between items and reorders the queue. Also checks that a model trained on
the solver's own counts predicts subtree sizes of larger boards."""
import numpy as np
import pytest

import nqm
from nqm import learned
from conftest import golden_json

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_model(ctx):
    X, Y = learned.build_dataset(ctx, [(11, 3), (11, 4), (12, 3), (12, 4), (12, 5), (13, 4), (13, 5)])
    m, info = learned.train(ctx, X, Y, epochs=15, batch=2048, lr=3e-3)
    assert info["val_r2"] > 0.6, info
    yield m
    m.close()


def test_learned_split_counts_exact(ctx, small_model, known_totals, node_totals):
    for n in (14, 15, 16, 17):
        for kw in (dict(), dict(split_rounds=3), dict(flags=nqm.TUNE_NO_ORDER), dict(split_log2=4.0)):
            c, nd, _ = ctx.solve_batch([(n, ())], model=small_model, **kw)
            assert int(c[0]) == known_totals[n], (n, kw)
            assert int(nd[0]) == node_totals[n], (n, kw)


def test_learned_split_mixed_batch(ctx, small_model):
    g = golden_json("c3_batch.json")["solves"]
    solves = [(r["n"], tuple(r["prefix"])) for r in g]
    c, nd, _ = ctx.solve_batch(solves, model=small_model, depth=4)
    assert np.array_equal(c, np.array([r["count"] for r in g], np.uint64))
    assert np.array_equal(nd, np.array([r["nodes"] for r in g], np.uint64))


def test_learned_with_restart_budget(ctx, small_model, known_totals):
    c, _, st = ctx.solve_batch([(15, ()), (14, (2, 5))], model=small_model, budget=300)
    assert int(c[0]) == known_totals[15]
    assert st.launches > 1


def test_model_generalises_to_a_larger_board(ctx, small_model):
    # N=14 at depth 5 leaves 9 rows, a depth the training boards also had
    # (N=12 depth 3, N=13 depth 4); the board itself was never seen.
    # what largest-first ordering needs is the ranking of subtree sizes
    e = learned.prediction_error(ctx, small_model, 14, 5)
    assert e["spearman"] > 0.6, e
    assert e["top1pct_overlap"] > 0.2, e


def test_graph_with_model_is_refused(ctx, small_model):
    rc, *_ = ctx.solve_batch([(12, ())], raw=True, model=small_model, flags=nqm.TUNE_GRAPH)
    assert rc == nqm.EUNSUPPORTED
