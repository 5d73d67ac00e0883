"""Device pool generation (n_queens.cu:27-219 semantics) against the oracle:
byte-identical pools, the reference's 17 pool-file sizes, the lock-row search."""
import numpy as np
import pytest

import nqm
import oracle
from conftest import golden_json

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [4, 5, 8, 9, 12, 13, 16])
def test_pools_byte_identical(ctx, n):
    for lock in range(1, n):
        for half in (False, True):
            if oracle.pool_count(n, lock, mirror=half) > 3_000_000:
                continue
            want = oracle.pool_generate(n, lock, mirror=half)
            got = ctx.pool_generate(n, lock, nqm.MIRROR_WEIGHT if half else 0)
            assert got.shape == want.shape and np.array_equal(got, want), (n, lock, half)


def test_reference_pool_sizes(ctx):
    for e in golden_json("pool_sizes.json")["pools"]:
        assert ctx.pool_count(e["n"], e["lock"]) == e["count"], e


def test_per_column_counts(ctx):
    for n, lock in ((11, 5), (14, 6)):
        tot, pc = ctx.pool_count(n, lock, nqm.MIRROR_WEIGHT, per_col=True)
        wt, wpc = oracle.pool_count(n, lock, mirror=True, per_col=True)
        assert tot == wt and np.array_equal(pc, wpc[:n])


def test_choose_lock_matches_reference_search(ctx):
    assert ctx.choose_lock(18, 80_000_000) == (8, 37956796)
    for n, k in ((12, 200_000), (14, 1_000_000), (16, 5_000_000), (10, 5000)):
        assert ctx.choose_lock(n, k) == oracle.lock_depth(n, k, mirror=True), (n, k)


def test_generated_pool_counts(ctx, known_totals):
    buf, cnt = ctx.pool_generate_device(17, 7, nqm.MIRROR_WEIGHT)
    try:
        st = ctx.count_pool_device(17, 7, buf, cnt, flags=nqm.MIRROR_WEIGHT)
    finally:
        buf.free()
    assert st.total == known_totals[17]


def test_bad_pool_args(ctx):
    with pytest.raises(nqm.NqmError):
        ctx.pool_count(8, 8)
    with pytest.raises(nqm.NqmError):
        ctx.pool_count(8, 0)
    with pytest.raises(nqm.NqmError):
        ctx.choose_lock(8, 1)
