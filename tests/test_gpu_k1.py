"""Parity of the static-pool HIP kernel (K1) with the CPU oracle, through the
C ABI (nqm_count_pool / nqm_count_pool_device / nqm_solver_driver)."""
import numpy as np
import pytest

import nqm
import oracle
from conftest import golden_npz

pytestmark = pytest.mark.gpu


def _full_levels(n, d):
    return sum(oracle.level_counts(n, d)[1:d + 1])


def test_c1_n12_depth3_per_prefix(ctx, known_totals):
    g = golden_npz("n12_d3_prefixes.npz")
    assert ctx.count_pool(12, 3, g["pool"]) == known_totals[12]
    d = nqm.DeviceBuffer.from_host(ctx, g["pool"])
    try:
        st, ps, pn = ctx.count_pool_device(12, 3, d, len(g["pool"]), per_prefix=True)
    finally:
        d.free()
    assert np.array_equal(ps, g["sols"])
    assert np.array_equal(pn, g["nodes"])
    assert st.total == 14200 and st.nodes == 855310


def test_c2_n16_depth5_full_and_half(ctx, known_totals, node_totals):
    full = oracle.pool_generate(16, 5, mirror=False)
    assert full.shape[0] == 141812
    d = nqm.DeviceBuffer.from_host(ctx, full)
    try:
        st = ctx.count_pool_device(16, 5, d, full.shape[0])
    finally:
        d.free()
    assert st.total == known_totals[16]
    assert st.nodes + _full_levels(16, 5) == node_totals[16]
    half = oracle.pool_generate(16, 5, mirror=True)
    assert half.shape[0] == 70906
    assert ctx.count_pool(16, 5, half, flags=nqm.MIRROR_WEIGHT) == known_totals[16]


def test_golden_n16_sample(ctx):
    g = golden_npz("n16_d5_sample.npz")
    d = nqm.DeviceBuffer.from_host(ctx, g["pool"])
    try:
        _, ps, pn = ctx.count_pool_device(16, 5, d, len(g["pool"]), per_prefix=True)
    finally:
        d.free()
    assert np.array_equal(ps, g["sols"]) and np.array_equal(pn, g["nodes"])


@pytest.mark.parametrize("n", range(4, 18))
def test_known_totals_sweep(ctx, n, known_totals):
    d = max(1, min(n - 2, 3 + (n - 8) // 2 if n > 8 else 2))
    for mirror in (False, True):
        pool = oracle.pool_generate(n, d, mirror=mirror)
        flags = nqm.MIRROR_WEIGHT if mirror else 0
        assert ctx.count_pool(n, d, pool, flags=flags) == known_totals[n], (n, d, mirror)


@pytest.mark.parametrize("n", [8, 10, 11, 12])
def test_every_depth_per_prefix(ctx, n):
    for d in range(1, min(8, n)):
        pool = oracle.pool_generate(n, d, mirror=False)
        _, _, ws, wn = oracle.count_pool(n, d, pool, per_prefix=True)
        buf = nqm.DeviceBuffer.from_host(ctx, pool)
        try:
            _, gs, gn = ctx.count_pool_device(n, d, buf, pool.shape[0], per_prefix=True)
        finally:
            buf.free()
        assert np.array_equal(gs, ws), (n, d)
        assert np.array_equal(gn, wn), (n, d)


def test_random_prefixes_mixed_depths(ctx):
    rng = np.random.default_rng(7)
    for n in range(8, 17):
        for d in (1, 2, 3, 4, 5, 6):
            if d >= n:
                continue
            full = oracle.pool_generate(n, d, mirror=False)
            pick = full[rng.choice(full.shape[0], min(300, full.shape[0]), replace=False)]
            tot, nd, ws, wn = oracle.count_pool(n, d, pick, per_prefix=True)
            buf = nqm.DeviceBuffer.from_host(ctx, pick)
            try:
                st, gs, gn = ctx.count_pool_device(n, d, buf, pick.shape[0], per_prefix=True)
            finally:
                buf.free()
            assert np.array_equal(gs, ws) and np.array_equal(gn, wn), (n, d)
            assert st.total == tot and st.nodes == nd


def test_unpruned_and_full_prefixes(ctx):
    # prefixes whose next row has no free cell, and complete boards
    n = 8
    pool = oracle.pool_generate(n, 7, mirror=False)
    _, _, ws, wn = oracle.count_pool(n, 7, pool, per_prefix=True)
    buf = nqm.DeviceBuffer.from_host(ctx, pool)
    try:
        _, gs, gn = ctx.count_pool_device(n, 7, buf, pool.shape[0], per_prefix=True)
    finally:
        buf.free()
    assert np.array_equal(gs, ws) and np.array_equal(gn, wn)
    # depth == n: one solution per complete board, no nodes below
    sols = []
    for p in oracle.pool_generate(n, 7, mirror=False):
        free = [c for c in range(n) if c not in p]
        q = [int(c) for c in p] + free
        if all(abs(q[i] - q[j]) != j - i for i in range(n) for j in range(i + 1, n)):
            sols.append(q)
    full = np.array(sols, np.uint8)
    assert ctx.count_pool(n, n, full) == 92


def test_depth_zero_is_a_full_solve(ctx, known_totals):
    for n in range(1, 13):
        assert ctx.count_pool(n, 0, np.zeros((1, 0), np.uint8), count=1) == known_totals[n]
    assert ctx.count_pool(10, 0, np.zeros((1, 0), np.uint8), count=3) == 3 * known_totals[10]


def test_invalid_prefix_is_rejected(ctx):
    bad = np.array([[0, 1, 5]], np.uint8)  # rows 0 and 1 attack diagonally
    rc, _ = ctx.count_pool_raw(8, 3, bad)
    assert rc == nqm.EPREFIX
    rc, _ = ctx.count_pool_raw(8, 2, np.array([[0, 9]], np.uint8))  # column >= n
    assert rc == nqm.EPREFIX
    # the context stays usable after an error
    assert ctx.count_pool(8, 1, np.arange(8, dtype=np.uint8).reshape(8, 1)) == 92


def test_bad_arguments(ctx):
    pool = np.zeros((1, 1), np.uint8)
    assert ctx.count_pool_raw(0, 1, pool)[0] == nqm.EINVAL
    assert ctx.count_pool_raw(32, 1, pool)[0] == nqm.EINVAL
    assert ctx.count_pool_raw(8, 9, np.zeros((1, 9), np.uint8))[0] == nqm.EINVAL
    assert ctx.count_pool_raw(8, 1, pool, count=0)[0] == nqm.EINVAL
    assert ctx.count_pool_raw(8, 1, pool, flags=0x80)[0] == nqm.EINVAL
    assert ctx.count_pool_raw(8, 0, np.zeros((1, 0), np.uint8), flags=nqm.MIRROR_WEIGHT,
                              count=1)[0] == nqm.EINVAL


def test_solver_driver_reference_layout(known_totals):
    for n, lock in ((12, 3), (14, 4)):
        pool = oracle.pool_generate(n, lock, mirror=True)
        states = nqm.states_from_pool(n, lock, pool)
        assert states.shape[1] == nqm.state_size(n)
        assert nqm.solver_driver(n, states, lock) == known_totals[n]


def test_n18_full_count_depth8(ctx, known_totals, node_totals):
    """N=18 on the reference's default pool (lock 8, mirror-half, 37,956,796
    prefixes): the count and the executed half-tree nodes."""
    pool = oracle.pool_generate(18, 8, mirror=True)
    assert pool.shape[0] == 37956796
    buf = nqm.DeviceBuffer.from_host(ctx, pool)
    try:
        st = ctx.count_pool_device(18, 8, buf, pool.shape[0], flags=nqm.MIRROR_WEIGHT)
    finally:
        buf.free()
    assert st.total == known_totals[18]
    # half tree: N=18 is even, so exactly half of every level is left-half
    half_levels = sum(oracle.level_counts(18, 8)[1:9]) // 2
    assert st.nodes + half_levels == node_totals[18] // 2
