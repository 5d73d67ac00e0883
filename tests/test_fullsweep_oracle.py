"""Pins the full DoubleSweep oracle (oracle/nq_fullsweep.c) on the CPU.

The reference describes full DoubleSweep (README.md:6, :22, :40-41) but does
not implement it, so the search is defined in nq_fullsweep.c. What pins the
oracle's restatement of that definition:
  * solutions: the reference's known-answer table (test.cu:18) and brute
    force over permutations;
  * solutions and placements: an independent restatement on an explicit board
    (fullsweep_ref.py: no bit masks), and a count derived by hand for N=4;
  * the recursion between a state's count and its children's, the thread
    split, and the pruning bound against the light search's full tree.
"""
import os
import sys

import pytest

import oracle
from conftest import golden_json

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import fullsweep_ref  # noqa: E402
from test_oracle import brute_count  # noqa: E402


@pytest.mark.parametrize("n", range(1, 11))
def test_matches_explicit_board_restatement(n):
    assert oracle.fullsweep_count(n, 1) == fullsweep_ref.fullsweep_count(n)


def test_hand_count_n4():
    # Row 1 (the centre row of n=4) at column 0, weight 2: placement 1. The
    # column pass finds column 1 free only in row 3 (placement 2), then
    # column 2 only in row 0 (3), column 3 only in row 2 (4): a solution,
    # weight 2. Row 1 at column 1: placement 5; row 0 is left with column 3
    # alone (6), after which row 2 has no free cell. 2 solutions, 6 placements.
    assert oracle.fullsweep_count(4, 1) == (2, 6)


def test_solutions_are_the_known_totals(known_totals):
    for n in range(1, 16):
        assert oracle.fullsweep_count(n, 8)[0] == known_totals[n], n


@pytest.mark.parametrize("n", [5, 6, 7, 8])
def test_solutions_by_brute_force(n):
    assert oracle.fullsweep_count(n, 1)[0] == brute_count(n)


def test_threads_agree():
    assert oracle.fullsweep_count(13, 1) == oracle.fullsweep_count(13, 8)


def test_unhalved_search_prunes_the_light_tree(known_totals, node_totals):
    # the search from the empty board without mirror halving: same solutions,
    # fewer placements than the light search's full tree (every one of its
    # placements is a placement of the light search too, at fewer of them)
    for n in range(8, 14):
        s, d = oracle.fullsweep_subtree(n, (0, 0, 0, 0))
        assert s == known_totals[n]
        assert d < node_totals[n], (n, d, node_totals[n])


def _place(n, st, r, c):
    P, C, D1, D2 = st
    return (P | 1 << r, C | 1 << c, D1 | 1 << (r + c), D2 | 1 << (c + n - 1 - r))


@pytest.mark.parametrize("n", [9, 11])
def test_subtree_is_propagation_plus_children(n):
    # count(state) = propagated placements + children + the children's counts
    m = (n - 1) // 2
    for c in range(n):
        st = _place(n, (0, 0, 0, 0), m, c)
        s, d = oracle.fullsweep_subtree(n, st)
        kids, prop = oracle.fullsweep_children(n, st)
        cs = [oracle.fullsweep_subtree(n, k) for k in kids]
        if kids:
            assert s == sum(x[0] for x in cs)
            assert d == prop + len(kids) + sum(x[1] for x in cs)
        # the mirror image of a root child holds as many solutions
        assert s == oracle.fullsweep_subtree(n, _place(n, (0, 0, 0, 0), m, n - 1 - c))[0]


def test_branch_rule_changes_the_placements():
    # the placement check is sensitive to the branch rule: a top-down rule
    # gives the same solutions and other placements
    s, d = oracle.fullsweep_count(9, 1)
    s2, d2 = fullsweep_ref.fullsweep_count(9, fullsweep_ref.top_down_row)
    assert s2 == s and d2 != d


def test_fixture_agrees_with_oracle(known_totals):
    rows = golden_json("fullsweep.json")["counts"]
    for n in range(1, 15):
        assert (rows[str(n)]["solutions"], rows[str(n)]["nodes"]) == oracle.fullsweep_count(n, 8)
    for k, v in rows.items():
        assert v["solutions"] == known_totals[int(k)]
