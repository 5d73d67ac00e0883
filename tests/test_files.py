"""The .nqsf pool file (nq_utils.cu:127-259) and the reference's nq_state_t
layout (n_queens.cuh:100-113): host-only entry points, no GPU needed.

No pool file written by the reference exists (the pre-generated files named
at test.cu:88-104 are not in the repository), so the byte layout is pinned by
the format comment at nq_utils.cu:127-144 through the oracle's writer, and
reading is checked on files built by hand from that comment."""
import struct

import numpy as np
import pytest

import nqm
import oracle


def test_write_matches_reference_layout(tmp_path):
    pool = oracle.pool_generate(10, 4, mirror=True)
    path = tmp_path / "p.nqsf"
    nqm.nqsf_write(str(path), 10, 4, pool, flags=nqm.NQSF_REFERENCE_FLAGS)
    assert path.read_bytes() == oracle.nqsf_bytes(10, 4, pool, flags=3)


def test_round_trip(tmp_path):
    for n, lock, half in ((8, 3, False), (12, 5, True), (16, 2, False)):
        pool = oracle.pool_generate(n, lock, mirror=half)
        path = tmp_path / f"p{n}.nqsf"
        nqm.nqsf_write(str(path), n, lock, pool)
        n2, l2, fl, p2 = nqm.nqsf_read(str(path), expect_n=n)
        assert (n2, l2, fl) == (n, lock, 3)
        assert np.array_equal(p2, pool)


def test_read_hand_built_multi_header_file(tmp_path):
    # header + two segments of the same row (the writer emits one per row;
    # a reader must accept any split)
    pre = [(0, 2), (0, 3), (1, 3)]
    b = struct.pack("<IQBBQ", 1, 3, 6, 2, 3)
    b += struct.pack("<BQ", 2, 2) + bytes([0, 2, 0, 3])
    b += struct.pack("<BQ", 2, 1) + bytes([1, 3])
    path = tmp_path / "h.nqsf"
    path.write_bytes(b)
    n, lock, fl, pool = nqm.nqsf_read(str(path))
    assert (n, lock, fl) == (6, 2, 3)
    assert [tuple(int(c) for c in p) for p in pool] == pre


@pytest.mark.parametrize("mutate,status", [
    (lambda b: struct.pack("<I", 2) + b[4:], nqm.EFORMAT),           # version != 1 (-3)
    (lambda b: b[:-1], nqm.EFORMAT),                                  # truncated
    (lambda b: b[:22] + struct.pack("<BQ", 3, 3) + b[31:], nqm.EFORMAT),  # wrong row
    (lambda b: b[:31] + bytes([9]) + b[32:], nqm.EFORMAT),            # column >= n
])
def test_malformed_files_are_rejected(tmp_path, mutate, status):
    pool = np.array([[0, 2], [0, 3], [1, 3]], np.uint8)
    b = oracle.nqsf_bytes(6, 2, pool, flags=3)
    path = tmp_path / "m.nqsf"
    path.write_bytes(mutate(b))
    rc, _ = nqm.nqsf_read(str(path), raw=True)
    assert rc == status


def test_flag_and_n_checks(tmp_path):
    pool = oracle.pool_generate(8, 3, mirror=False)
    path = tmp_path / "f.nqsf"
    nqm.nqsf_write(str(path), 8, 3, pool, flags=nqm.NQSF_FLAG_EXPERIMENTAL)
    rc, _ = nqm.nqsf_read(str(path), expect_flags=3, raw=True)
    assert rc == nqm.EFLAGS  # util_read_nq_states_from_stream returns -4 here
    rc, _ = nqm.nqsf_read(str(path), expect_n=9, raw=True)
    assert rc == nqm.EINVAL
    rc, _ = nqm.nqsf_read(str(tmp_path / "missing.nqsf"), raw=True)
    assert rc == nqm.EIO


def test_state_layout_round_trip():
    for n, lock in ((12, 3), (19, 5), (20, 4)):
        pool = oracle.pool_generate(n, lock, mirror=True)[:500]
        st = nqm.pool_to_states(n, lock, pool)
        assert st.shape[1] == nqm.state_size(n)
        # same bytes as the Python restatement of place_queen_at
        assert np.array_equal(st, nqm.states_from_pool(n, lock, pool))
        assert np.array_equal(nqm.states_to_pool(n, st, lock), pool)


def test_state_fields_match_place_queen_at():
    st = nqm.pool_to_states(8, 2, np.array([[1, 3]], np.uint8))[0]
    diag = int.from_bytes(st[0:8].tobytes(), "little")
    anti = int.from_bytes(st[8:16].tobytes(), "little")
    cols = int.from_bytes(st[16:20].tobytes(), "little")
    # queen (0,1): diag bit 1, anti bit 1+8; queen (1,3): diag bit 4, anti bit 3+7
    assert diag == (1 << 1) | (1 << 4)
    assert anti == (1 << 9) | (1 << 10)
    assert cols == 0b1010 and st[20] == 2
    assert list(st[21:29]) == [1, 3] + [0xFF] * 6
