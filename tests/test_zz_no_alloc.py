"""The allocation-free count (nqm_count_pool_device_static) against the
allocating one (nqm_count_pool_device) on the same pools, through the C ABI:
three calls in a row, each smaller than the one before, so that every call
reuses the scratch the previous one left behind. Counts are exact (bit-equal
solutions and node numbers, and the known totals of test.cu:18)."""
import numpy as np
import pytest

import nqm
import oracle

pytestmark = pytest.mark.gpu


def _static(ctx, n, depth, d_pool, count, flags, res):
    ctx.count_pool_device_static(n, depth, d_pool, count, res, flags=flags)
    return [int(v) for v in res.to_host(np.uint64, 3)]


def test_static_equals_allocating_count(ctx, known_totals):
    res = nqm.DeviceBuffer(ctx, 24)
    try:
        # the reference's default N=18 pool (lock 8, 37,956,796 half-pool prefixes)
        big, count = ctx.pool_generate_device(18, 8, nqm.MIRROR_WEIGHT)
        try:
            st = ctx.count_pool_device(18, 8, big, count, flags=nqm.MIRROR_WEIGHT)
            got = _static(ctx, 18, 8, big, count, nqm.MIRROR_WEIGHT, res)
        finally:
            big.free()
        assert st.total == known_totals[18]
        assert got == [st.total, st.nodes, 0]
        # smaller inputs next: N=16 depth 5 half pool, then N=12 depth 3 full pool
        for n, depth, mirror in ((16, 5, True), (12, 3, False)):
            pool = oracle.pool_generate(n, depth, mirror=mirror)
            flags = nqm.MIRROR_WEIGHT if mirror else 0
            d = nqm.DeviceBuffer.from_host(ctx, pool)
            try:
                st = ctx.count_pool_device(n, depth, d, pool.shape[0], flags=flags)
                got = _static(ctx, n, depth, d, pool.shape[0], flags, res)
            finally:
                d.free()
            assert st.total == known_totals[n]
            assert got == [st.total, st.nodes, 0], (n, depth)
    finally:
        res.free()


def test_static_flags_invalid_prefix_and_rejects_oversize(ctx):
    res = nqm.DeviceBuffer(ctx, 24)
    d = nqm.DeviceBuffer.from_host(ctx, np.array([[0, 0], [0, 2]], np.uint8))  # row 1 repeats column 0
    try:
        got = _static(ctx, 8, 2, d, 2, 0, res)
        assert got[2] != 0
        # 2^26 + 1 prefixes need one workgroup more than the scratch holds:
        # refused on the host before anything is launched
        with pytest.raises(nqm.NqmError) as e:
            ctx.count_pool_device_static(8, 2, d, (1 << 26) + 1, res)
        assert e.value.status == nqm.ERANGE
    finally:
        d.free()
        res.free()
