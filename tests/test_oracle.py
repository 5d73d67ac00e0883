"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Every check here is independent of the oracle's own formula: the reference's
known-answer table (test.cu:18), its pool-file names (test.cu:88-104), brute
force over permutations, the mirror symmetry, and the node recursion. A
last group plants plausible mistakes in a pure-Python copy of the counting
loop and shows that these checks catch them.
"""
import itertools

import numpy as np
import pytest

import oracle
from conftest import golden_json, golden_npz


def brute_count(n, prefix=()):
    """Solutions extending `prefix` by brute force over all permutations."""
    k = len(prefix)
    rest = [c for c in range(n) if c not in prefix]
    cnt = 0
    for perm in itertools.permutations(rest):
        q = list(prefix) + list(perm)
        if all(abs(q[i] - q[j]) != j - i for i in range(n) for j in range(i + 1, n)):
            cnt += 1
    return cnt


def brute_levels(n):
    """Valid partial placements of k queens in rows 0..k-1, k=0..n."""
    lv = [1] + [0] * n
    for k in range(1, n + 1):
        for q in itertools.permutations(range(n), k):
            if all(abs(q[i] - q[j]) != j - i for i in range(k) for j in range(i + 1, k)):
                lv[k] += 1
    return lv


# ---- totals -----------------------------------------------------------------

def test_known_totals_tiny_boards(known_totals):
    for n in (1, 2, 3):
        s, _ = oracle.count_prefix(n, [])
        assert s == known_totals[n]


@pytest.mark.parametrize("n", range(4, 15))
def test_known_totals_from_pools(n, known_totals):
    d = min(3, n - 1)
    for mirror in (False, True):
        pool = oracle.pool_generate(n, d, mirror=mirror)
        tot, _ = oracle.count_pool(n, d, pool, mirror=mirror)
        assert tot == known_totals[n]


def test_known_totals_n15_n16(known_totals):
    for n, d in ((15, 4), (16, 5)):
        pool = oracle.pool_generate(n, d, mirror=True)
        tot, _ = oracle.count_pool(n, d, pool, mirror=True)
        assert tot == known_totals[n]


@pytest.mark.parametrize("n", range(4, 9))
def test_brute_force_totals(n):
    s, _ = oracle.count_prefix(n, [])
    assert s == brute_count(n)


def test_brute_force_per_prefix_n7():
    pool = oracle.pool_generate(7, 2, mirror=False)
    _, _, ps, _ = oracle.count_pool(7, 2, pool, per_prefix=True)
    for pre, s in zip(pool, ps):
        assert int(s) == brute_count(7, tuple(int(c) for c in pre))


# ---- nodes ------------------------------------------------------------------

@pytest.mark.parametrize("n", range(4, 8))
def test_level_counts_brute_force(n):
    assert oracle.level_counts(n) == brute_levels(n)


def test_node_totals_match_fixture(node_totals):
    for n in range(8, 14):
        assert sum(oracle.level_counts(n)[1:]) == node_totals[n]


@pytest.mark.parametrize("n,d", [(8, 3), (10, 4), (12, 3), (13, 5)])
def test_pool_nodes_plus_prefix_levels_is_full_tree(n, d, node_totals):
    pool = oracle.pool_generate(n, d, mirror=False)
    _, nd = oracle.count_pool(n, d, pool)
    lv = oracle.level_counts(n)
    # Pruned prefixes (row d has no free cell) have no nodes below them.
    assert nd + sum(lv[1:d + 1]) == node_totals[n]


def test_node_recursion_n10():
    """nodes(p) = sum over children c of (1 + nodes(c))."""
    pool = oracle.pool_generate(10, 3, mirror=False)
    for pre in pool[:40]:
        pre = [int(c) for c in pre]
        _, nd = oracle.count_prefix(10, pre)
        s = 0
        for c in range(10):
            try:
                _, k = oracle.count_prefix(10, pre + [c])
            except ValueError:
                continue
            s += 1 + k
        assert nd == s


def test_full_prefix_counts_one_or_zero():
    pool = oracle.pool_generate(8, 7, mirror=False)
    for pre in pool:
        pre = [int(c) for c in pre]
        free = [c for c in range(8) if c not in pre]
        s, nd = oracle.count_prefix(8, pre)
        ok = all(abs(free[0] - pre[r]) != 7 - r for r in range(7))
        assert (s, nd) == ((1, 1) if ok else (0, 0))


# ---- mirror -----------------------------------------------------------------

@pytest.mark.parametrize("n", [9, 10, 11])
def test_mirror_symmetry_per_prefix(n):
    pool = oracle.pool_generate(n, 3, mirror=False)
    _, _, ps, pn = oracle.count_pool(n, 3, pool, per_prefix=True)
    lookup = {tuple(int(c) for c in p): (int(s), int(k)) for p, s, k in zip(pool, ps, pn)}
    for p, v in lookup.items():
        m = tuple(n - 1 - c for c in p)
        assert lookup[m] == v


@pytest.mark.parametrize("n", range(5, 14))
def test_mirror_closed_form(n, known_totals):
    _, per = oracle.pool_count(n, 1, mirror=False, per_col=True), None
    t = []
    for c in range(n):
        s, _ = oracle.count_prefix(n, [c])
        t.append(s)
    closed = 2 * sum(t[: n // 2]) + (t[n // 2] if n % 2 else 0)
    assert closed == known_totals[n]


# ---- pool generation (n_queens.cu:27-219) -----------------------------------

def test_pool_sizes_match_reference_file_names():
    for e in golden_json("pool_sizes.json")["pools"]:
        assert oracle.pool_count(e["n"], e["lock"], mirror=False) == e["count"], e


def test_pool_half_sizes():
    assert oracle.pool_count(12, 3, mirror=False) == 756
    assert oracle.pool_count(16, 5, mirror=False) == 141812
    assert oracle.pool_count(16, 5, mirror=True) == 70906
    # even n: the half pool is exactly half the full one
    assert oracle.pool_count(18, 6, mirror=True) * 2 == oracle.pool_count(18, 6, mirror=False)


def test_pool_equals_brute_force_enumeration():
    n, lock = 8, 4
    pool = oracle.pool_generate(n, lock, mirror=False)
    want = []
    for q in itertools.permutations(range(n), lock):
        if not all(abs(q[i] - q[j]) != j - i for i in range(lock) for j in range(i + 1, lock)):
            continue
        free = [c for c in range(n) if c not in q and all(abs(c - q[r]) != lock - r for r in range(lock))]
        if free:
            want.append(q)
    want.sort()
    got = [tuple(int(c) for c in p) for p in pool]
    assert got == want  # same set, lexicographic order


def test_mirror_pool_is_left_half_prefix_of_full():
    full = oracle.pool_generate(11, 4, mirror=False)
    half = oracle.pool_generate(11, 4, mirror=True)
    assert np.array_equal(full[full[:, 0] < 6], half)


def test_lock_depth_replay():
    # n_queens.cu:135-192 replayed for the default request of 8e7 states
    # (NQ_solver.cu:427-430): N=18 half pools 1.20M (6), 7.41M (7),
    # 37.96M (8), 158.8M (9 > 9.6e7) -> lock 8.
    assert oracle.lock_depth(18, 80_000_000, mirror=True) == (8, 37956796)
    assert oracle.pool_count(18, 7, mirror=True) < 37956796
    # A small request: the chosen pool fits the cap 1.2*k and the next lock's
    # pool either exceeds the cap or shrinks (downslope bound).
    k = 200_000
    lk, cnt = oracle.lock_depth(12, k, mirror=True)
    assert cnt == oracle.pool_count(12, lk, mirror=True) and cnt <= int(k * 1.2)
    nxt = oracle.pool_count(12, lk + 1, mirror=True)
    assert nxt > int(k * 1.2) or nxt < cnt


def test_lock_depth_failure_when_first_pool_too_big():
    # floor(ln 10 / ln 8) = 1: the lock-1 half pool has 4 states <= 12, then
    # lock 2 ... the search never fails; a request of 1 gives lock 0 -> fail.
    with pytest.raises(ValueError):
        oracle.lock_depth(8, 1, mirror=True)


# ---- golden vectors ---------------------------------------------------------

def test_golden_n12_d3_table(known_totals):
    g = golden_npz("n12_d3_prefixes.npz")
    pool = oracle.pool_generate(12, 3, mirror=False)
    assert np.array_equal(pool, g["pool"])
    _, _, ps, pn = oracle.count_pool(12, 3, pool, per_prefix=True)
    assert np.array_equal(ps, g["sols"]) and np.array_equal(pn, g["nodes"])
    assert int(g["sols"].sum()) == known_totals[12]
    assert int(g["nodes"].sum()) + 12 + 110 + 756 == 856188  # levels 1..3 of N=12


# ---- .nqsf layout (nq_utils.cu:127-193) --------------------------------------

def test_nqsf_layout():
    pool = oracle.pool_generate(8, 3, mirror=True)
    b = oracle.nqsf_bytes(8, 3, pool, flags=3)
    assert int.from_bytes(b[0:4], "little") == 1
    assert int.from_bytes(b[4:12], "little") == 3
    assert b[12] == 8 and b[13] == 3
    assert int.from_bytes(b[14:22], "little") == len(pool)
    assert b[22] == 3 and int.from_bytes(b[23:31], "little") == len(pool)
    assert b[31:] == pool.tobytes()


# ---- planted mistakes: the checks above must catch each -----------------------

def _py_count(n, prefix, mistake=None):
    """Pure-Python copy of the counting loop with optional planted mistakes."""
    M = (1 << n) - 1
    UNSET = 0xFF
    q = [UNSET] * n
    cols = diag = anti = 0

    def free(r):
        return ~(cols | ((diag >> r) | (anti >> (n - r))) & 0xFFFFFFFF) & M

    def place(r, c):
        nonlocal cols, diag, anti
        cols |= 1 << c
        q[r] = c
        diag |= (1 << c) << (r + (1 if mistake == "diag_shift" else 0))
        anti |= (1 << c) << (n - r)

    def remove(r, c):
        nonlocal cols, diag, anti
        cols &= ~(1 << c)
        q[r] = UNSET
        diag &= ~((1 << c) << (r + (1 if mistake == "diag_shift" else 0)))
        anti &= ~((1 << c) << (n - r))

    for r, c in enumerate(prefix):
        place(r, c)
    lock = len(prefix)
    row = lock
    cnt = 0
    while row >= lock:
        while row >= lock:
            qi = q[row]
            f = free(row)
            if qi != UNSET:
                f &= M << (qi + 1)
                remove(row, qi)
            if not f:
                row -= 1
            else:
                place(row, (f & -f).bit_length() - 1)
                if row < n - 1:
                    row += 1
                break
        while row >= 0 and q[row] == UNSET:
            f = free(row)
            if bin(f).count("1") == 1:
                col = (f & -f).bit_length() - 1
                if mistake == "ffs_plus_one":  # the inactive branch, GPU_NQueens.cu:401
                    col = min(col + 2, n - 1)
                place(row, col)
                if row < n - 1:
                    row += 1
            else:
                break
        cnt += cols == M
    return cnt


def test_python_copy_matches_oracle():
    for n in (6, 7, 8):
        assert _py_count(n, [0]) == oracle.count_prefix(n, [0])[0]


@pytest.mark.parametrize("mistake", ["diag_shift", "ffs_plus_one"])
def test_planted_mistakes_are_caught(mistake, known_totals):
    caught = False
    for n in (6, 7, 8):
        tot = sum(_py_count(n, [c], mistake) for c in range(n))
        caught |= tot != known_totals[n]
    assert caught


def test_planted_mirror_mistake_is_caught(known_totals):
    # weighting x2 for c0 <= n/2 instead of c0 < n/2 (wrong for odd n)
    n = 9
    t = [oracle.count_prefix(n, [c])[0] for c in range((n + 1) // 2)]
    wrong = sum(2 * s for c, s in enumerate(t) if c <= n // 2)
    right = sum((2 if c < n // 2 else 1) * s for c, s in enumerate(t))
    assert right == known_totals[n] and wrong != known_totals[n]


# ---- symmetry-reduced counting (the reference's future work, README.md:38-45)

def test_symmetry_reduction_matches_brute_force():
    # every solution's 8 images, a class counted at its smallest member; the
    # brute force also checks that the classes' orbit sizes add up to the total
    for n in range(1, 11):
        assert oracle.sym_count(n)[:2] == oracle.sym_brute(n), n


def test_symmetry_reduction_published_counts(known_totals):
    unique = {int(k): v for k, v in golden_json("unique_totals.json")["unique"].items()}
    for n in range(1, 16):
        u, t, nodes = oracle.sym_count(n)
        assert u == unique[n], n
        assert t == known_totals[n], n


def test_symmetry_reduction_visits_less_than_the_mirror_halved_tree(node_totals):
    # the point of the method: under half of the mirror-halved placements
    for n in (12, 14):
        _, _, nodes = oracle.sym_count(n)
        assert nodes < 0.5 * node_totals[n] / 2, (n, nodes)
