"""Parity of the persistent work-queue path (K2: expand + drain with
in-wave donation, restart budget, HIP graph) with the oracle and the
reference's known answers, through nqm_solve_batch."""
import numpy as np
import pytest

import nqm
import oracle
from conftest import golden_json

pytestmark = pytest.mark.gpu


def _full_nodes(n):
    return sum(oracle.level_counts(n)[1:])


@pytest.mark.parametrize("n", range(1, 16))
def test_empty_boards_counts_and_nodes(ctx, n, known_totals):
    want_nodes = _full_nodes(n)
    for flags in (0, nqm.TUNE_NO_MIRROR):
        c, nd, st = ctx.solve_batch([(n, ())], flags=flags)
        assert int(c[0]) == known_totals[n], (n, flags)
        assert int(nd[0]) == want_nodes, (n, flags)


def test_n16_n17_every_depth(ctx, known_totals, node_totals):
    for n in (16, 17):
        for depth in (3, 4, 5, 6, 7):
            c, nd, _ = ctx.solve_batch([(n, ())], depth=depth)
            assert int(c[0]) == known_totals[n] and int(nd[0]) == node_totals[n], (n, depth)


def test_c3_mixed_batch_512(ctx):
    g = golden_json("c3_batch.json")["solves"]
    solves = [(r["n"], tuple(r["prefix"])) for r in g]
    want_c = np.array([r["count"] for r in g], np.uint64)
    want_n = np.array([r["nodes"] for r in g], np.uint64)
    c, nd, st = ctx.solve_batch(solves)
    assert np.array_equal(c, want_c)
    assert np.array_equal(nd, want_n)
    assert st.total == int(want_c.sum())


@pytest.mark.parametrize("tune", [
    dict(depth=3), dict(depth=5), dict(depth=7),
    dict(waves_per_cu=4), dict(waves_per_cu=8), dict(waves_per_cu=12), dict(waves_per_cu=16),
    dict(threads=64), dict(threads=128), dict(threads=256), dict(refill_idle=1), dict(refill_idle=64),
    dict(threads=256, budget=300), dict(refill_idle=1, budget=40),
    dict(flags=nqm.TUNE_NO_DONATE), dict(flags=nqm.TUNE_NO_MIRROR), dict(stack_rows=6),
    dict(flags=nqm.TUNE_GRAPH),
])
def test_c3_batch_every_tuning(ctx, tune):
    g = golden_json("c3_batch.json")["solves"]
    solves = [(r["n"], tuple(r["prefix"])) for r in g]
    c, nd, _ = ctx.solve_batch(solves, **tune)
    assert np.array_equal(c, np.array([r["count"] for r in g], np.uint64)), tune
    assert np.array_equal(nd, np.array([r["nodes"] for r in g], np.uint64)), tune


@pytest.mark.parametrize("budget", [1, 7, 64, 1000, 50000])
def test_restart_budget_is_exact(ctx, budget):
    g = golden_json("c3_batch.json")["solves"][:96]
    solves = [(r["n"], tuple(r["prefix"])) for r in g]
    c, nd, st = ctx.solve_batch(solves, budget=budget)
    assert np.array_equal(c, np.array([r["count"] for r in g], np.uint64)), budget
    assert np.array_equal(nd, np.array([r["nodes"] for r in g], np.uint64)), budget
    if budget <= 64:
        assert st.launches > 1  # restarts happened
    # rounds never multiply the work: a round parks at most the piece and its
    # split-off bottom row of each lane, and hands back at most what its
    # waves held
    assert st.parked_max <= 2 * st.lanes, (budget, st.parked_max, st.lanes)
    assert st.handed_back_max <= st.hand_back_room, (budget, st.handed_back_max, st.hand_back_room)


def test_restart_single_big_solve(ctx, known_totals, node_totals):
    c, nd, st = ctx.solve_batch([(16, ())], budget=2000, depth=4)
    assert int(c[0]) == known_totals[16] and int(nd[0]) == node_totals[16]
    assert st.launches > 2
    assert st.rounds > 2 and st.parked_max <= 2 * st.lanes
    assert st.handed_back_max <= st.hand_back_room


def test_batch_alone_vs_in_mix(ctx):
    g = golden_json("c3_batch.json")["solves"]
    solves = [(r["n"], tuple(r["prefix"])) for r in g]
    mixed, mixed_nodes, _ = ctx.solve_batch(solves)
    for i in (0, 17, 255, 511):
        c, nd, _ = ctx.solve_batch([solves[i]])
        assert c[0] == mixed[i] and nd[0] == mixed_nodes[i]


def test_graph_replay_is_stable(ctx, known_totals):
    for _ in range(3):
        c, _, _ = ctx.solve_batch([(15, ()), (14, (1, 3)), (12, ())], flags=nqm.TUNE_GRAPH)
        assert int(c[0]) == known_totals[15] and int(c[2]) == known_totals[12]
        assert int(c[1]) == oracle.count_prefix(14, [1, 3])[0]


def test_graph_plan_replays_and_is_dropped(ctx, known_totals):
    """A graph-mode call that repeats the previous one launches the kept
    graph again; any other call (another batch, or a plain one) drops it.
    Every replay is checked against the eager result, with eager work and
    synchronizations between the replays."""
    a = [(15, ()), (14, (1, 3)), (12, ())]
    b = [(13, ()), (11, (0,))]
    eager = {}
    for name, solves in (("a", a), ("b", b)):
        c, nd, _ = ctx.solve_batch(solves)
        eager[name] = (list(c), list(nd))
    assert eager["a"][0][0] == known_totals[15] and eager["b"][0][0] == known_totals[13]
    prev = None
    for name in ("a", "a", "a", "b", "a", "b", "b", "plain-a", "a", "a"):
        solves = a if name.endswith("a") else b
        flags = 0 if name.startswith("plain") else nqm.TUNE_GRAPH
        c, nd, st = ctx.solve_batch(solves, flags=flags)
        key = "a" if name.endswith("a") else "b"
        assert (list(c), list(nd)) == eager[key], name
        assert st.total == sum(eager[key][0])
        # a replay is one launch; building the graph also ran the planning levels
        assert (st.launches == 1) == (name == prev), (name, prev, st.launches)
        prev = name


def test_full_and_dead_prefixes(ctx):
    full = (0, 4, 7, 5, 2, 6, 1, 3)  # an 8-queens solution
    dead = (0, 2, 4, 1, 3)           # row 5 of N=8 has no free cell? check with oracle
    c, nd, _ = ctx.solve_batch([(8, full), (8, dead), (1, ()), (2, ()), (3, ())])
    assert int(c[0]) == 1 and int(nd[0]) == 0
    want = oracle.count_prefix(8, list(dead))
    assert (int(c[1]), int(nd[1])) == want
    assert list(c[2:]) == [1, 0, 0]


def test_batch_rejects_invalid(ctx):
    rc, *_ = ctx.solve_batch([(8, (0, 1))], raw=True)
    assert rc == nqm.EPREFIX
    rc, *_ = ctx.solve_batch([(0, ())], raw=True)
    assert rc == nqm.EINVAL
    rc, *_ = ctx.solve_batch([(8, ())], raw=True, flags=0x100)
    assert rc == nqm.EINVAL


def test_n18_persistent(ctx, known_totals, node_totals):
    c, nd, st = ctx.solve_batch([(18, ())])
    assert int(c[0]) == known_totals[18]
    assert int(nd[0]) == node_totals[18]
    assert st.nodes == node_totals[18] // 2  # mirror-halved executed nodes (even n)
