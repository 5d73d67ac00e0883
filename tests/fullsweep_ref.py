"""Full DoubleSweep search restated on an explicit board (test infrastructure).

An independent restatement of the search oracle/nq_fullsweep.c defines (the
reference's future work, README.md:6, :22, :40-41; centre-out row order of
the unbuilt kernel, GPU_NQueens.cu:207 and :214-336): the board is a list of
queen columns per row and every free-cell test walks the queens, with no
bitsets or diagonal masks, so that a wrong shift or diagonal index in the C
oracle (or in the device kernel) shows up as a different count.

Semantics, as in the oracle:
  propagate: repeat {rows ascending, then columns ascending; an open line
    with no free cell kills the board, one with a single free cell gets its
    queen (a placement)} until a row pass and the column pass after it place
    nothing;
  branch on the open row nearest m = (n-1)//2 (m, m+1, m-1, ...), its free
    columns ascending;
  the root branches on row m with left-half columns weighing 2 and the middle
  column of odd n weighing 1.
Returns (solutions, placements).
"""


def _attacked(board, r, c):
    for rr, cc in enumerate(board):
        if cc is None:
            continue
        if cc == c or rr == r or abs(rr - r) == abs(cc - c):
            return True
    return False


def _row_free(board, n, r):
    return [c for c in range(n) if not _attacked(board, r, c)]


def _col_free(board, n, c):
    return [r for r in range(n) if board[r] is None and not _attacked(board, r, c)]


def _propagate(board, n, stats):
    while True:
        placed = False
        for r in [r for r in range(n) if board[r] is None]:
            f = _row_free(board, n, r)
            if not f:
                return False
            if len(f) == 1:
                board[r] = f[0]
                stats[1] += 1
                placed = True
        used = {c for c in board if c is not None}
        for c in [c for c in range(n) if c not in used]:
            g = _col_free(board, n, c)
            if not g:
                return False
            if len(g) == 1:
                board[g[0]] = c
                stats[1] += 1
                placed = True
        if not placed:
            return True


def _branch_row(board, n):
    m = (n - 1) // 2
    for d in range(n):
        for r in ((m + d, m - d) if d else (m,)):
            if 0 <= r < n and board[r] is None:
                return r
    raise AssertionError("no open row")


def _node(board, n, w, stats, order):
    if not _propagate(board, n, stats):
        return
    if all(c is not None for c in board):
        stats[0] += w
        return
    r = order(board, n)
    for c in _row_free(board, n, r):
        child = list(board)
        child[r] = c
        stats[1] += 1
        _node(child, n, w, stats, order)


def fullsweep_count(n, order=_branch_row):
    """(solutions, placements) of the mirror-halved full-sweep search.
    `order` picks the branch row (the default is the centre-out rule)."""
    stats = [0, 0]
    m = (n - 1) // 2
    for c in range(n // 2 + (n & 1)):
        board = [None] * n
        board[m] = c
        stats[1] += 1
        _node(board, n, 2 if c < n // 2 else 1, stats, order)
    return stats[0], stats[1]


def top_down_row(board, n):
    """A different branch rule (first open row), for sensitivity checks."""
    return next(r for r in range(n) if board[r] is None)
