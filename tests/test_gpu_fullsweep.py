"""Full DoubleSweep counting on the device (K5; the reference's future work,
README.md:6, :22, :40-41) against the full-sweep oracle (oracle/nq_fullsweep.c)
and the known totals (test.cu:18): solutions and the search's own placements
must equal the oracle's at every n, whatever level the expansion stops at."""
import pytest

import nqm
import oracle
from conftest import golden_json

pytestmark = pytest.mark.gpu


def test_matches_oracle_small(ctx):
    for n in range(1, 15):
        t, st = ctx.count_fullsweep(n)
        assert (t, st.nodes) == oracle.fullsweep_count(n, 8), n


@pytest.mark.parametrize("target", [1, 7, 64, 4096, 1 << 16, 1 << 22])
def test_any_expansion_target_gives_the_same_counts(ctx, target):
    # small targets hand the drain states near the root (deep stacks); large
    # ones leave it the last rows (or nothing)
    for n in (11, 13, 14):
        t, st = ctx.count_fullsweep(n, target)
        assert (t, st.nodes) == oracle.fullsweep_count(n, 8), (n, target)


def test_n15_to_n18_against_fixture(ctx, known_totals):
    rows = golden_json("fullsweep.json")["counts"]
    for n in (15, 16, 17, 18):
        t, st = ctx.count_fullsweep(n)
        assert t == known_totals[n], n
        assert st.nodes == rows[str(n)]["nodes"], n
        assert st.search_ms > 0 and st.kernel_ms >= st.search_ms


def test_repeated_calls_agree(ctx):
    first = ctx.count_fullsweep(16)
    for _ in range(2):
        t, st = ctx.count_fullsweep(16)
        assert (t, st.nodes) == (first[0], first[1].nodes)


def test_limits(ctx):
    with pytest.raises(nqm.NqmError) as e:
        ctx.count_fullsweep(22)
    assert e.value.status == nqm.EUNSUPPORTED
    with pytest.raises(nqm.NqmError) as e:
        ctx.count_fullsweep(0)
    assert e.value.status == nqm.EINVAL
    # the context still counts after a refused call
    assert ctx.count_fullsweep(12)[0] == 14200
