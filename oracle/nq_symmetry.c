/* nq_symmetry.c -- CPU oracle for symmetry-reduced counting (test
 * infrastructure only).
 *
 * The reference lists "transformation elimination" as future work
 * (README.md:38-45) and has no implementation; this file restates the
 * published corner / non-corner method for counting solution classes under
 * the 8 symmetries of the board, as the checker for a later device version.
 *
 *   - Corner case: a queen at (0,0). No solution with a corner queen is
 *     symmetric (a diagonal mirror would put two queens on one anti-diagonal),
 *     so each class counts 8. Of the two class members with a queen at (0,0),
 *     the one whose row-1 queen column b is smaller than the row of its
 *     column-1 queen is kept: rows 2..b-1 may not use column 1.
 *   - Non-corner case: row 0's queen at column b1 < n-1-b1 (= b2). The class
 *     member whose top queen is nearest a corner is kept: rows above b1 and
 *     below b2 avoid the side columns (so both are used by row b2), and the
 *     last row avoids the columns within b1 of a corner. A full board is then
 *     compared with its 90, 180 and 270 degree rotations: it is kept only if
 *     no rotation is smaller (row by row), and it stands for 2, 4 or 8
 *     solutions as it equals its 90 or 180 degree rotation or neither.
 *
 * Pinned by orc_sym_brute (every solution's 8 images, canonical = smallest)
 * and by the published unique-solution counts (OEIS A002562). */
#include <stdint.h>
#include <string.h>

#include "nq_oracle.h"

typedef struct {
    unsigned n, last;  /* board size, last row (n - 1) */
    uint32_t mask, top, side, lastmask, endbit;
    unsigned b1, b2;
    uint32_t board[32]; /* one-hot queen column per row */
    uint64_t c2, c4, c8, nodes;
} orc_sym;

/* Counts the full board if no rotation of it is smaller: in c2, c4 or c8 as
 * it equals its 90 degree rotation, its 180 degree rotation, or neither. */
static void orc_sym_check(orc_sym* s) {
    const uint32_t* b = s->board;
    const int last = (int)s->last;
    /* 90 degrees: possible only if row b2's queen is in column 0 */
    if (b[s->b2] == 1u) {
        int own;
        uint32_t ptn = 2u;
        for (own = 1; own <= last; ++own, ptn <<= 1) {
            uint32_t bit = 1u;
            for (int you = last; b[you] != ptn && b[own] >= bit; --you) bit <<= 1;
            if (b[own] > bit) return;
            if (b[own] < bit) break;
        }
        if (own > last) {
            s->c2++;
            return;
        }
    }
    /* 180 degrees: possible only if the last row's queen mirrors row 0's */
    if (b[last] == s->endbit) {
        int own, you;
        for (own = 1, you = last - 1; own <= last; ++own, --you) {
            uint32_t bit = 1u;
            for (uint32_t ptn = s->top; ptn != b[you] && b[own] >= bit; ptn >>= 1) bit <<= 1;
            if (b[own] > bit) return;
            if (b[own] < bit) break;
        }
        if (own > last) {
            s->c4++;
            return;
        }
    }
    /* 270 degrees: possible only if row b1's queen is in the last column */
    if (b[s->b1] == s->top) {
        int own;
        uint32_t ptn = s->top >> 1;
        for (own = 1; own <= last; ++own, ptn >>= 1) {
            uint32_t bit = 1u;
            for (int you = 0; b[you] != ptn && b[own] >= bit; ++you) bit <<= 1;
            if (b[own] > bit) return;
            if (b[own] < bit) break;
        }
    }
    s->c8++;
}

/* corner case below row 1 (row 0's queen at column 0, row 1's at b1) */
static void orc_sym_corner(orc_sym* s, unsigned y, uint32_t left, uint32_t down, uint32_t right) {
    uint32_t f = s->mask & ~(left | down | right);
    if (y == s->last) {
        s->nodes += (uint64_t)__builtin_popcount(f);
        if (f) s->c8++;
        return;
    }
    if (y < s->b1) f &= ~2u;
    s->nodes += (uint64_t)__builtin_popcount(f);
    while (f) {
        const uint32_t bit = f & (0u - f);
        f ^= bit;
        s->board[y] = bit;
        orc_sym_corner(s, y + 1, (left | bit) << 1, down | bit, (right | bit) >> 1);
    }
}

/* non-corner case below row 0 (row 0's queen at column b1) */
static void orc_sym_side(orc_sym* s, unsigned y, uint32_t left, uint32_t down, uint32_t right) {
    uint32_t f = s->mask & ~(left | down | right);
    if (y == s->last) {
        if (f && !(f & s->lastmask)) {
            s->nodes += 1;
            s->board[y] = f;
            orc_sym_check(s);
        }
        return;
    }
    if (y < s->b1) {
        f &= ~s->side;
    } else if (y == s->b2) {
        if (!(down & s->side)) return;               /* a side column would be left to the rows below */
        if ((down & s->side) != s->side) f &= s->side; /* this row takes the other side column */
    }
    s->nodes += (uint64_t)__builtin_popcount(f);
    while (f) {
        const uint32_t bit = f & (0u - f);
        f ^= bit;
        s->board[y] = bit;
        orc_sym_side(s, y + 1, (left | bit) << 1, down | bit, (right | bit) >> 1);
    }
}

int orc_sym_count(unsigned n, uint64_t* unique, uint64_t* total, uint64_t* nodes) {
    if (n < 1 || n > 31) return -1;
    if (n < 4) {  /* too small for the bound rows: the brute force is exact */
        if (nodes) *nodes = 0;
        return orc_sym_brute(n, unique, total);
    }
    orc_sym s;
    memset(&s, 0, sizeof(s));
    s.n = n;
    s.last = n - 1;
    s.mask = (1u << n) - 1u;
    s.top = 1u << (n - 1);
    /* corner case */
    s.board[0] = 1u;
    for (s.b1 = 2; s.b1 < s.last; ++s.b1) {
        const uint32_t bit = 1u << s.b1;
        s.board[1] = bit;
        s.nodes += 2;  /* the two placements of rows 0 and 1 */
        orc_sym_corner(&s, 2, (2u | bit) << 1, 1u | bit, bit >> 1);
    }
    /* non-corner case */
    s.side = s.lastmask = s.top | 1u;
    s.endbit = s.top >> 1;
    for (s.b1 = 1, s.b2 = s.last - 1; s.b1 < s.b2; ++s.b1, --s.b2) {
        const uint32_t bit = 1u << s.b1;
        s.board[0] = bit;
        s.nodes += 1;
        orc_sym_side(&s, 1, bit << 1, bit, bit >> 1);
        s.lastmask |= (s.lastmask >> 1) | (s.lastmask << 1);
        s.endbit >>= 1;
    }
    if (unique) *unique = s.c2 + s.c4 + s.c8;
    if (total) *total = 2 * s.c2 + 4 * s.c4 + 8 * s.c8;
    if (nodes) *nodes = s.nodes;
    return 0;
}

/* ---- brute force: every solution, its 8 images, canonical = smallest ---- */

typedef struct {
    unsigned n;
    uint8_t q[32];
    uint64_t unique, total, orbit_sum;
} orc_brute;

/* image k of board q (row -> column) under the 8 symmetries */
static void orc_image(const uint8_t* q, unsigned n, int k, uint8_t* out) {
    for (unsigned r = 0; r < n; ++r) {
        const unsigned c = q[r];
        unsigned rr = r, cc = c;
        switch (k) {
            case 0: rr = r; cc = c; break;                  /* identity */
            case 1: rr = c; cc = n - 1 - r; break;          /* rotate 90 */
            case 2: rr = n - 1 - r; cc = n - 1 - c; break;  /* rotate 180 */
            case 3: rr = n - 1 - c; cc = r; break;          /* rotate 270 */
            case 4: rr = r; cc = n - 1 - c; break;          /* mirror columns */
            case 5: rr = n - 1 - r; cc = c; break;          /* mirror rows */
            case 6: rr = c; cc = r; break;                  /* transpose */
            default: rr = n - 1 - c; cc = n - 1 - r; break; /* anti-transpose */
        }
        out[rr] = (uint8_t)cc;
    }
}

static void orc_brute_leaf(orc_brute* b) {
    uint8_t img[8][32];
    int smallest = 1, distinct = 0;
    for (int k = 0; k < 8; ++k) {
        orc_image(b->q, b->n, k, img[k]);
        if (memcmp(img[k], b->q, b->n) < 0) smallest = 0;
        int seen = 0;
        for (int j = 0; j < k && !seen; ++j) seen = memcmp(img[j], img[k], b->n) == 0;
        distinct += !seen;
    }
    b->total++;
    if (smallest) {
        b->unique++;
        b->orbit_sum += (uint64_t)distinct;
    }
}

static void orc_brute_rec(orc_brute* b, unsigned r, uint32_t cols, uint32_t ld, uint32_t rd) {
    if (r == b->n) {
        orc_brute_leaf(b);
        return;
    }
    const uint32_t mask = (1u << b->n) - 1u;
    uint32_t f = mask & ~(cols | ld | rd);
    while (f) {
        const uint32_t bit = f & (0u - f);
        f ^= bit;
        b->q[r] = (uint8_t)__builtin_ctz(bit);
        orc_brute_rec(b, r + 1, cols | bit, ((ld | bit) << 1) & mask, (rd | bit) >> 1);
    }
}

int orc_sym_brute(unsigned n, uint64_t* unique, uint64_t* total) {
    if (n < 1 || n > 16) return -1;
    orc_brute b;
    memset(&b, 0, sizeof(b));
    b.n = n;
    orc_brute_rec(&b, 0, 0, 0, 0);
    /* each class is counted once, by a member whose orbit it sums */
    if (b.orbit_sum != b.total) return -2;
    if (unique) *unique = b.unique;
    if (total) *total = b.total;
    return 0;
}
