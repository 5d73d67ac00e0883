// sym_rules.h -- the rules of the symmetry-reduced search, shared by the
// thread-per-item kernel (sym.hip), the persistent drain (k4_sym.hip) and
// their host code.
//
// The board's 8 symmetries are removed by the published corner / non-corner
// method for bit-board N-Queens (Kenji Takahashi's symmetry elimination):
//   - corner case: a queen at (0,0); rows 2..b1-1 avoid column 1, where b1 is
//     row 1's column; no such solution is symmetric, so each counts 8;
//   - non-corner case: row 0's queen at column b1 < b2 = n-1-b1; rows above
//     b1 avoid the side columns, row b2 completes them, the last row avoids
//     the columns within b1 of a corner, and a full board counts only if no
//     rotation of it is smaller (2, 4 or 8 as it equals its 90 or 180 degree
//     rotation or neither).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nqm {

struct SymCase {
    uint32_t n, last, mask, top, side, lastmask, endbit, b1, b2, corner;
};

__host__ __device__ inline SymCase sym_case(uint32_t n, uint32_t b1, uint32_t corner) {
    SymCase c;
    c.n = n;
    c.last = n - 1u;
    c.mask = (1u << n) - 1u;
    c.top = 1u << (n - 1u);
    c.side = c.top | 1u;
    c.b1 = b1;
    c.b2 = c.last - b1;
    c.corner = corner;
    const uint32_t low = (1u << b1) - 1u;  // columns within b1 of a corner
    c.lastmask = low | (low << (n - b1));
    c.endbit = c.top >> b1;
    return c;
}

// Candidates of row y under the case's rules (0: the row ends the branch).
// The rules are selects, not branches (rows y < b1, y == b2 and y == last
// are distinct: b1 < b2 < last), so lanes at different rows stay together.
__host__ __device__ inline uint32_t sym_free(const SymCase& c, uint32_t y, uint32_t l, uint32_t d, uint32_t r) {
    uint32_t f = c.mask & ~(l | d | r);
    // rows above b1: the corner case avoids column 1, the other the side columns
    f &= (y < c.b1) ? ~(c.corner ? 2u : c.side) : ~0u;
    if (!c.corner) {
        // row b2: a side column left to the rows below ends the branch, one
        // side column taken makes this row take the other
        const uint32_t ds = d & c.side;
        const uint32_t at_b2 = ds == 0u ? 0u : (ds != c.side ? (f & c.side) : f);
        f = y == c.b2 ? at_b2 : f;
        // the last row avoids the columns within b1 of a corner
        f = (y == c.last && (f & c.lastmask) != 0u) ? 0u : f;
    }
    return f;
}

// 2, 4 or 8 if the full board b (one-hot column per row) is the smallest of
// its rotations, as it equals its 90 or 180 degree rotation or neither; 0 if
// a rotation is smaller.
template <class Board>
__host__ __device__ inline uint32_t sym_check(const SymCase& c, const Board& b) {
    const int last = (int)c.last;
    if (b(c.b2) == 1u) {  // 90 degrees
        int own;
        uint32_t ptn = 2u;
        for (own = 1; own <= last; ++own, ptn <<= 1) {
            uint32_t bit = 1u;
            for (int you = last; b(you) != ptn && b(own) >= bit; --you) bit <<= 1;
            if (b(own) > bit) return 0u;
            if (b(own) < bit) break;
        }
        if (own > last) return 2u;
    }
    if (b(last) == c.endbit) {  // 180 degrees
        int own, you;
        for (own = 1, you = last - 1; own <= last; ++own, --you) {
            uint32_t bit = 1u;
            for (uint32_t ptn = c.top; ptn != b(you) && b(own) >= bit; ptn >>= 1) bit <<= 1;
            if (b(own) > bit) return 0u;
            if (b(own) < bit) break;
        }
        if (own > last) return 4u;
    }
    if (b(c.b1) == c.top) {  // 270 degrees
        uint32_t ptn = c.top >> 1;
        for (int own = 1; own <= last; ++own, ptn >>= 1) {
            uint32_t bit = 1u;
            for (int you = 0; b(you) != ptn && b(own) >= bit; ++you) bit <<= 1;
            if (b(own) > bit) return 0u;
            if (b(own) < bit) break;
        }
    }
    return 8u;
}

// sym_check on a board packed as 5-bit columns (row r at bit 5 r of lo for
// r < 12, at bit 5 (r - 12) of hi above), with the same result. Each
// rotation's row is read from the inverse permutation (column -> row)
// instead of searched for: row i of the 90-degree rotation has its queen in
// column last - row(column i), of the 180-degree one in column
// last - column(last - i), of the 270-degree one in column row(column
// last - i). The comparisons stop where the loops of sym_check stop; the
// inverse is built only for a board that reaches the 90 or 270 degree test.
__host__ __device__ inline uint32_t sym_check_packed(const SymCase& c, uint64_t lo, uint64_t hi) {
    auto field = [](uint64_t a, uint64_t b, uint32_t k) -> uint32_t {
        return (uint32_t)((k < 12u ? a >> (5u * k) : b >> (5u * (k - 12u))) & 31u);
    };
    const uint32_t last = c.last;
    uint64_t ilo = 0, ihi = 0;  // row of each column, packed the same way
    if (field(lo, hi, c.b2) == 0u || field(lo, hi, c.b1) == last) {
        for (uint32_t r = 0; r <= last; ++r) {
            const uint32_t k = field(lo, hi, r);
            if (k < 12u)
                ilo |= (uint64_t)r << (5u * k);
            else
                ihi |= (uint64_t)r << (5u * (k - 12u));
        }
    }
    if (field(lo, hi, c.b2) == 0u) {  // 90 degrees
        uint32_t own = 1;
        for (; own <= last; ++own) {
            const uint32_t a = field(lo, hi, own), r = last - field(ilo, ihi, own);
            if (a > r) return 0u;
            if (a < r) break;
        }
        if (own > last) return 2u;
    }
    if (field(lo, hi, last) == c.b2) {  // 180 degrees (endbit is column b2)
        uint32_t own = 1;
        for (; own <= last; ++own) {
            const uint32_t a = field(lo, hi, own), r = last - field(lo, hi, last - own);
            if (a > r) return 0u;
            if (a < r) break;
        }
        if (own > last) return 4u;
    }
    if (field(lo, hi, c.b1) == last) {  // 270 degrees
        for (uint32_t own = 1; own <= last; ++own) {
            const uint32_t a = field(lo, hi, own), r = field(ilo, ihi, last - own);
            if (a > r) return 0u;
            if (a < r) break;
        }
    }
    return 8u;
}

}  // namespace nqm
