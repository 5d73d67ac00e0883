"""Serving by continuous batching: every request's count (and node count) is
the same in the mix as alone, and equals the oracle's, whatever the admission
limits and round budget."""
import pytest

import nqm
import oracle
from nqm import serving, workloads

pytestmark = pytest.mark.gpu


def _stream(count=150, n_lo=6, n_hi=14, seed=99):
    return workloads.serving_stream(seed=seed, count=count, n_lo=n_lo, n_hi=n_hi, rate=4000.0)


def _want(n, prefix):
    return oracle.count_prefix(n, list(prefix))


def _assert_bounded_rounds(rs):
    # rounds never multiply the work: a round parks at most the piece and its
    # split-off bottom row of each lane, and hands back at most the queue
    # items its waves held
    assert rs["parked_max"] <= 2 * rs["lanes"], rs
    assert rs["handed_back_max"] <= rs["hand_back_room"], rs


@pytest.mark.parametrize("cfg", [
    dict(), dict(max_active=4), dict(budget=64), dict(queue_bytes=64 * 1024, max_active=8),
    dict(depth=2, budget=300), dict(flags=nqm.TUNE_NO_MIRROR, max_active=3, budget=1000),
])
def test_stream_counts_match_oracle(ctx, cfg):
    reqs = _stream()
    srv = ctx.server(**cfg)
    try:
        res, _ = serving.replay(srv, reqs)
        rs = srv.round_stats()
    finally:
        srv.close()
    assert len(res) == len(reqs)
    _assert_bounded_rounds(rs)
    for tag, (_, n, pre) in enumerate(reqs):
        c, nd = _want(n, pre)
        assert (res[tag][0], res[tag][1]) == (c, nd), (tag, n, pre, cfg)


def test_mix_equals_alone(ctx):
    reqs = _stream(count=60, n_lo=10, n_hi=15, seed=5)
    srv = ctx.server(max_active=16, budget=2000)
    try:
        res, _ = serving.replay(srv, reqs)
    finally:
        srv.close()
    for tag, (_, n, pre) in enumerate(reqs):
        c, nd, _ = ctx.solve_batch([(n, tuple(pre))])
        assert res[tag][0] == int(c[0]) and res[tag][1] == int(nd[0]), (tag, n, pre)


def test_overlapping_requests_share_rounds(ctx, known_totals):
    # a long request (many short rounds) and short ones that arrive while it runs
    reqs = [(0.0, 17, ())] + [(0.0002 * i, 8 + i % 4, ()) for i in range(1, 40)]
    srv = ctx.server(budget=200)
    try:
        res, _ = serving.replay(srv, reqs)
        rs = srv.round_stats()
    finally:
        srv.close()
    _assert_bounded_rounds(rs)
    assert rs["rounds"] > 10
    assert res[0][0] == known_totals[17]
    for tag in range(1, 40):
        assert res[tag][0] == known_totals[8 + tag % 4]
    # short requests finished while the long one was still in flight
    assert min(res[t][5] for t in range(1, 40)) < res[0][5]
    assert res[0][2] > 1


def test_rejects_invalid_request(ctx):
    srv = ctx.server()
    try:
        with pytest.raises(nqm.NqmError):
            srv.submit([(8, (0, 1))], [0])
    finally:
        srv.close()


def test_serving_config_slice_with_large_boards(ctx, known_totals):
    # The serving workload's own stream (seed, N 10..20 with weight 1/N^2,
    # 30% with a 1-3 queen prefix): its first 200 requests, arriving 100 per
    # second, on the default server configuration. They include N=18, 19 and
    # 20 boards, which the server stages deeper (row n-1-11 for n > 18).
    reqs = workloads.serving_stream(count=200, rate=100.0)
    sizes = {n for _, n, _ in reqs}
    assert {18, 19, 20} <= sizes
    srv = ctx.server()
    try:
        res, _ = serving.replay(srv, reqs)
        rs = srv.round_stats()
    finally:
        srv.close()
    assert len(res) == len(reqs)
    for tag, (_, n, pre) in enumerate(reqs):
        if not pre:  # an empty board: the known total (test.cu:18)
            assert res[tag][0] == known_totals[n], (tag, n)
        else:  # a prefixed request: the same request alone
            c, nd, _ = ctx.solve_batch([(n, tuple(pre))])
            assert (res[tag][0], res[tag][1]) == (int(c[0]), int(nd[0])), (tag, n, pre)
    _assert_bounded_rounds(rs)
