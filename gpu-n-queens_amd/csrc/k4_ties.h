// k4_ties.h -- the search behind a tie record of the symmetric drain (K4),
// shared by k4_research_kernel and host-side checks.
//
// A record {item, L, R} stands for the boards of the item's subtree whose
// last row has ld == L and rd == R (the drain's relative diagonals at row
// n-1) and that the drain records as ties. The search finds them again
// bottom-up: rows last-1, last-2, ..., r0 (the item's row), each row's
// column drawn from the columns that
//   * set a bit of L on the last row: c + (last - y) in L,
//   * set a bit of R there, or fall below bit 0 by then: c < last - y or
//     c - (last - y) in R,
//   * no queen of the item's prefix attacks (columns, and its diagonals
//     carried down from row r0),
//   * no queen already placed below attacks (U, UL, UR: columns and the two
//     diagonals, moved up one row per row),
// (at row r0 that is the item's own candidates: no rule of the non-corner
// case applies at row r0, which lies between b1 and b2). The last row takes
// the one column left. The drain's own board is one such board, so the search finds
// it; every other board it completes is checked against what the record
// says:
//   * the last queen is not attacked: not in L nor in R (the last row's
//     attacked columns),
//   * rd: the prefix's rd bits plus each placed queen's rd bit (those of
//     columns c >= last - y) equal R (ld needs no check: each placed queen
//     sets a distinct bit of L, so the bits add up to L),
//   * row b2's rule of the non-corner case (sym_free): the side columns taken
//     above row b2 are not none, and unless both, row b2 takes the other,
//   * the last row avoids the columns within b1 of a corner,
//   * the board is a tie: its last queen on endb, or bit b1 in L.
// Bottom-up the rows with few candidates come first: about 18 placements per
// record at N=18 (measured on the host on a sample), against 28 top-down.
//
// The search runs in passes, as the drain does: a pass whose row is spent
// first goes back to the last row with candidates left (one stack entry per
// such row, not per row), then places the next candidate. So every pass
// places a queen, and lanes of a wave run the same few instructions.
// k4_research_kernel runs this pass in asm (tie_phase in k4_sym.hip), on the
// same stack entries, with the path kept as a shift register.
#pragma once

#include <cstdint>

#include "sym_rules.h"

namespace nqm {

// Stack entries a search holds at most: rows r0+1 .. last-1 (r0 is the
// drain's item row, ceil(n/2) - 1; 9 for n <= 20).
constexpr uint32_t kTieRows = 9;

struct TieSearch {
    // the record and its item
    uint32_t last, mask, L, R, r0, b2, side, bad, tie_l, endb;
    uint32_t cols0, ld0, rd0, rdk0;  // row r0: columns taken above it, its diagonals; rd0 at the last row
    // the search: current row y and its candidates left f; columns U and
    // diagonals UL, UR of the rows placed below y, as seen from row y; RP =
    // their rd bits on the last row; Ub2 / cb2 = columns of rows b2.. and
    // row b2's column (set when row b2 is placed); path = 5-bit columns of
    // rows r0.. at bit 5 (row - r0); sp = stack entries
    uint32_t y, f, U, UL, UR, RP, Ub2, cb2, sp;
    uint64_t path;

    __host__ __device__ inline uint32_t allowed(uint32_t yy) const {
        const uint32_t k = last - yy, j = yy - r0;
        return (L >> k) & ((R << k) | ((1u << k) - 1u)) & ~(cols0 | (ld0 << j) | (rd0 >> j)) & mask;
    }

    // n, the record's ld and rd (rd without the drain's marker), and the
    // item's row, b1, top_b1 flag, columns and diagonals
    __host__ __device__ inline void start(uint32_t n_, uint32_t L_, uint32_t R_, uint32_t row, uint32_t b1,
                                          bool top_b1, uint32_t cols, uint32_t ld, uint32_t rd) {
        const SymCase c = sym_case(n_, b1, 0u);
        last = c.last;
        mask = c.mask;
        L = L_;
        R = R_;
        r0 = row;
        b2 = c.b2;
        side = c.side;
        bad = L_ | R_ | c.lastmask;
        tie_l = L_ & (1u << b1);
        endb = (b1 < row && !top_b1) ? c.endbit : 0xffffffffu;
        cols0 = cols;
        ld0 = ld;
        rd0 = rd;
        rdk0 = rd >> (c.last - row);
        y = c.last - 1u;
        U = UL = UR = RP = Ub2 = cb2 = sp = 0u;
        path = 0;
        f = allowed(y);
    }

    // One pass. Returns 1 when the search is over, 2 with a board found
    // (*board: 5-bit columns of rows r0..last, row r0 at bit 0), else 0.
    // s4(i), s1(i): stack entry i, {f | y << 24, U, UL, UR} and RP (i <
    // kTieRows).
    template <class S4, class S1>
    __host__ __device__ inline int pass(S4 s4, S1 s1, uint64_t* board) {
        if (f == 0u) {
            if (sp == 0u) return 1;
            --sp;
            const uint4 e = s4(sp);
            f = e.x & 0xffffffu;
            y = e.x >> 24;
            U = e.y;
            UL = e.z;
            UR = e.w;
            RP = s1(sp);
        }
        const uint32_t bit = f & (0u - f);
        f ^= bit;
        const uint32_t sh = 5u * (y - r0);
        path = (path & ~(31ull << sh)) | ((uint64_t)__builtin_ctz(bit) << sh);
        if (y == b2) {
            Ub2 = U | bit;
            cb2 = bit;
        }
        const uint32_t k = last - y;
        if (y == r0) {  // every row placed but the last: its column is the one left
            const uint32_t all = U | bit, left = mask & ~(cols0 | all);
            const uint32_t rc = rdk0 | RP | (bit >> k);
            const uint32_t s = side & (cols0 | (all & ~Ub2));  // side columns above row b2
            const bool ok = left != 0u && rc == R && (left & bad) == 0u && s != 0u &&
                            (s == side || (cb2 & side) != 0u) && ((left & endb) | tie_l) != 0u;
            if (ok) {
                *board = path | ((uint64_t)__builtin_ctz(left) << (5u * k));
                return 2;
            }
            return 0;
        }
        if (f != 0u) {
            s4(sp) = make_uint4(f | (y << 24), U, UL, UR);
            s1(sp) = RP;
            ++sp;
        }
        U |= bit;
        UL = (UL | bit) >> 1;
        UR = (UR | bit) << 1;
        RP |= bit >> k;
        --y;
        f = allowed(y) & ~(U | UL | UR);
        return 0;
    }
};

}  // namespace nqm
