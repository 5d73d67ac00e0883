"""Symmetry-reduced counting on the device (the reference's future work,
README.md:38-45) against the symmetry oracle and the published counts:
unique classes (OEIS A002562), totals (test.cu:18), and the reduced search's
own placements, which must equal the oracle's. Both kernels: one thread per
host item (NQM_SYM_THREAD) and the persistent drain (NQM_SYM_DRAIN, whose
row rules ride in marker bits and whose tie boards are searched again)."""
import pytest

import nqm
import oracle
from conftest import golden_json

pytestmark = pytest.mark.gpu


def test_symmetric_matches_oracle(ctx):
    for n in range(1, 17):
        u, t, st = ctx.count_symmetric(n, nqm.SYM_THREAD)
        ou, ot, onodes = oracle.sym_count(n)
        assert (u, t) == (ou, ot), n
        if n >= 4:
            assert st.nodes == onodes, n


def test_symmetric_n17_n18(ctx, known_totals):
    unique = {int(k): v for k, v in golden_json("unique_totals.json")["unique"].items()}
    for n in (17, 18):
        u, t, st = ctx.count_symmetric(n, nqm.SYM_THREAD)
        assert u == unique[n] and t == known_totals[n], (n, u, t)


@pytest.mark.parametrize("n", range(10, 17))
def test_drain_matches_oracle_and_thread_kernel(ctx, n):
    ou, ot, onodes = oracle.sym_count(n)
    u, t, st = ctx.count_symmetric(n, nqm.SYM_DRAIN)
    assert (u, t, st.nodes) == (ou, ot, onodes), n
    u3, t3, st3 = ctx.count_symmetric(n, nqm.SYM_THREAD)
    assert (u3, t3, st3.nodes) == (u, t, st.nodes), n


def test_drain_is_the_default_from_n10(ctx):
    for n in (9, 12):
        u, t, st = ctx.count_symmetric(n)
        assert (u, t, st.nodes) == oracle.sym_count(n), n


def test_drain_n17_n18(ctx, known_totals):
    unique = {int(k): v for k, v in golden_json("unique_totals.json")["unique"].items()}
    for n in (17, 18):
        u, t, st = ctx.count_symmetric(n, nqm.SYM_DRAIN)
        assert u == unique[n] and t == known_totals[n], (n, u, t)
    # repeated calls: the same counts and placements
    u2, t2, st2 = ctx.count_symmetric(18, nqm.SYM_DRAIN)
    assert (u2, t2, st2.nodes) == (u, t, st.nodes)


def test_drain_n19(ctx, known_totals):
    unique = {int(k): v for k, v in golden_json("unique_totals.json")["unique"].items()}
    u, t, st = ctx.count_symmetric(19, nqm.SYM_DRAIN)
    assert u == unique[19] and t == known_totals[19], (u, t)


def test_drain_n20_the_largest_it_takes(known_totals):
    # ten stack rows (12 waves per CU) and nine tie rows; a first count sizes
    # its record chunks from an estimate and, when they run out, again from
    # the grabs the drain made. Its own context: the buffers of n = 20's tie
    # records (tens of GB) go when it closes.
    unique = {int(k): v for k, v in golden_json("unique_totals.json")["unique"].items()}
    print("symmetric drain n=20", flush=True)
    with nqm.Context(0) as big:
        u, t, st = big.count_symmetric(20, nqm.SYM_DRAIN)
    assert u == unique[20] and t == known_totals[20], (u, t)


def test_drain_rejects_sizes_outside_its_range(ctx):
    for n in (9, 21, 22):
        rc = ctx.lib.nqm_count_symmetric_path(ctx.handle, n, nqm.SYM_DRAIN, None, None, None)
        assert rc == nqm.EUNSUPPORTED, n


def test_drain_repeated_and_alternating_sizes(ctx):
    # a second count of an n takes the tie stage's sizes from the first (no
    # read-back between the drain and the tie stage); another n in between
    # resets them
    want = {n: oracle.sym_count(n)[:2] for n in (14, 15)}
    for n in (14, 14, 15, 14, 15, 15):
        u, t, _ = ctx.count_symmetric(n, nqm.SYM_DRAIN)
        assert (u, t) == want[n], n
