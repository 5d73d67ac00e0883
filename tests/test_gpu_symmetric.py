"""Symmetry-reduced counting on the device (the reference's future work,
README.md:38-45) against the symmetry oracle and the published counts:
unique classes (OEIS A002562), totals (test.cu:18), and the reduced search's
own placements, which must equal the oracle's."""
import pytest

import oracle
from conftest import golden_json

pytestmark = pytest.mark.gpu


def test_symmetric_matches_oracle(ctx):
    for n in range(1, 17):
        u, t, st = ctx.count_symmetric(n)
        ou, ot, onodes = oracle.sym_count(n)
        assert (u, t) == (ou, ot), n
        if n >= 4:
            assert st.nodes == onodes, n


def test_symmetric_n17_n18(ctx, known_totals):
    unique = {int(k): v for k, v in golden_json("unique_totals.json")["unique"].items()}
    for n in (17, 18):
        u, t, st = ctx.count_symmetric(n)
        assert u == unique[n] and t == known_totals[n], (n, u, t)
