"""A pool that went through a .nqsf file, counted on the device: the
reference's load-then-solve path (NQ_solver.cu:400-424 -> solve(),
NQ_solver.cu:192-209) with this library's reader (layout nq_utils.cu:127-259)
and its drop-in driver. Totals are the known answers of test.cu:18."""
import struct

import numpy as np
import pytest

import nqm
import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,lock", [(14, 4), (16, 5)])
def test_half_pool_through_file_and_driver(ctx, tmp_path, known_totals, n, lock):
    pool = ctx.pool_generate(n, lock, nqm.MIRROR_WEIGHT)  # the reference's half pool, on the device
    path = tmp_path / f"n{n}.nqsf"
    nqm.nqsf_write(str(path), n, lock, pool)  # flags 3: the reference's default build
    n2, l2, flags, back = nqm.nqsf_read(str(path), expect_n=n, expect_flags=nqm.NQSF_REFERENCE_FLAGS)
    assert (n2, l2, flags) == (n, lock, nqm.NQSF_REFERENCE_FLAGS)
    assert np.array_equal(back, pool)
    # the reference's state array (nq_state_t), then its driver's arguments
    states = nqm.pool_to_states(n, lock, back)
    assert np.array_equal(states, nqm.states_from_pool(n, lock, back))
    assert nqm.solver_driver(n, states, lock) == known_totals[n]
    assert ctx.count_pool(n, lock, back, flags=nqm.MIRROR_WEIGHT) == known_totals[n]


def test_hand_built_multi_segment_file(ctx, tmp_path, known_totals):
    # util_write_nq_states_to_stream (nq_utils.cu:145-193) writes a header and
    # one segment per populated row; a file whose row-3 states come in three
    # segments must read as the same pool
    pool = oracle.pool_generate(12, 3, mirror=True)
    cuts = [0, 100, 250, pool.shape[0]]
    b = struct.pack("<IQBBQ", 1, nqm.NQSF_REFERENCE_FLAGS, 12, 3, pool.shape[0])
    for lo, hi in zip(cuts, cuts[1:]):
        b += struct.pack("<BQ", 3, hi - lo) + pool[lo:hi].tobytes()
    path = tmp_path / "multi.nqsf"
    path.write_bytes(b)
    n, lock, flags, back = nqm.nqsf_read(str(path), expect_n=12)
    assert (n, lock, flags) == (12, 3, nqm.NQSF_REFERENCE_FLAGS)
    assert np.array_equal(back, pool)
    assert nqm.solver_driver(12, nqm.pool_to_states(12, 3, back), 3) == known_totals[12]
