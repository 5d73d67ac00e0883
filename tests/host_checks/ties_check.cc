// Host check of TieSearch (gpu-n-queens_amd/csrc/k4_ties.h), the search
// behind the symmetric drain's tie records, on small boards (its stack held
// to kTieRows entries).
//
// For every non-corner item at the drain's row R (the host's restatement of
// the item enumeration: row 0's column b1 < n-1-b1, rows below under
// sym_free), it walks the item's subtree as the drain does, takes every tie
// board (the last queen on endb, or bit b1 in the last row's ld) and groups
// them by record {item, ld, rd of the last row}. Then, for every record,
// TieSearch must find exactly the boards of its group, and the rotation
// check of what it finds must add up to the check of the boards themselves.
// usage: ties_check n_lo n_hi
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <map>
#include <set>
#include <tuple>
#include <vector>

#include "k4_ties.h"

using namespace nqm;

namespace {

struct Item {
    uint32_t f, cols, ld, rd, row, b1, top_b1;
    uint32_t col[32];  // columns of rows 0..row-1
};

uint32_t rows_done(uint32_t n) { return (n & 1u) ? (n - 1u) / 2u : n / 2u - 1u; }

// the board's columns packed 5 bits a row (rows < 12 in lo), as sym_check_packed reads them
void pack(const uint32_t* col, uint32_t n, uint64_t& lo, uint64_t& hi) {
    lo = hi = 0;
    for (uint32_t r = 0; r < n; ++r) {
        if (r < 12)
            lo |= (uint64_t)col[r] << (5u * r);
        else
            hi |= (uint64_t)col[r] << (5u * (r - 12u));
    }
}

int check(uint32_t n) {
    const uint32_t last = n - 1u, mask = (1u << n) - 1u, R = rows_done(n);
    std::vector<Item> items;
    for (uint32_t b1 = 1; 2u * b1 < last; ++b1) {
        const SymCase c = sym_case(n, b1, 0u);
        Item it{};
        it.col[0] = b1;
        std::function<void(uint32_t, uint32_t, uint32_t, uint32_t, uint32_t)> rec =
            [&](uint32_t y, uint32_t l, uint32_t d, uint32_t r, uint32_t top) {
                const uint32_t f = sym_free(c, y, l, d, r);
                if (!f) return;
                if (y == R) {
                    Item x = it;
                    x.f = f;
                    x.cols = d;
                    x.ld = l & mask;
                    x.rd = r;
                    x.row = y;
                    x.b1 = b1;
                    x.top_b1 = top;
                    items.push_back(x);
                    return;
                }
                for (uint32_t g = f; g; g &= g - 1u) {
                    const uint32_t bit = g & (0u - g);
                    it.col[y] = (uint32_t)__builtin_ctz(bit);
                    rec(y + 1u, ((l | bit) << 1) & mask, d | bit, (r | bit) >> 1,
                        top | (y == b1 && bit == c.top ? 1u : 0u));
                }
            };
        rec(1u, (1u << b1) << 1 & mask, 1u << b1, (1u << b1) >> 1, 0u);
    }
    uint64_t records = 0, boards = 0, passes = 0, want_classes = 0, want_weight = 0, got_classes = 0, got_weight = 0;
    for (const Item& it : items) {
        const SymCase c = sym_case(n, it.b1, 0u);
        const uint32_t endb = (it.b1 < it.row && !it.top_b1) ? c.endbit : 0xffffffffu, bitb1 = 1u << it.b1;
        std::map<std::pair<uint32_t, uint32_t>, std::set<std::pair<uint64_t, uint64_t>>> groups;
        uint32_t col[32];
        for (uint32_t r = 0; r < it.row; ++r) col[r] = it.col[r];
        std::function<void(uint32_t, uint32_t, uint32_t, uint32_t, uint32_t)> walk =
            [&](uint32_t y, uint32_t l, uint32_t d, uint32_t r, uint32_t f) {
                for (uint32_t g = f; g; g &= g - 1u) {
                    const uint32_t bit = g & (0u - g);
                    col[y] = (uint32_t)__builtin_ctz(bit);
                    if (y == last) {
                        if (((bit & endb) | (l & bitb1)) != 0u) {
                            uint64_t lo, hi;
                            pack(col, n, lo, hi);
                            groups[{l, r}].insert({lo, hi});
                            const uint32_t w = sym_check_packed(c, lo, hi);
                            want_classes += w != 0u;
                            want_weight += w;
                            ++boards;
                        }
                    } else {
                        const uint32_t l2 = (l | bit) << 1, d2 = d | bit, r2 = (r | bit) >> 1;
                        walk(y + 1u, l2, d2, r2, sym_free(c, y + 1u, l2, d2, r2));
                    }
                }
            };
        walk(it.row, it.ld, it.cols, it.rd, it.f);
        for (const auto& g : groups) {
            ++records;
            TieSearch ts;
            ts.start(n, g.first.first, g.first.second & mask, it.row, it.b1, it.top_b1 != 0u, it.cols, it.ld, it.rd);
            uint4 s4v[kTieRows] = {};
            uint32_t s1v[kTieRows] = {};
            auto s4 = [&](uint32_t i) -> uint4& {
                if (i >= kTieRows) {
                    std::printf("n=%u: stack entry %u\n", n, i);
                    std::exit(1);
                }
                return s4v[i];
            };
            auto s1 = [&](uint32_t i) -> uint32_t& { return s1v[i]; };
            std::set<std::pair<uint64_t, uint64_t>> found;
            for (int steps = 0;; ++steps) {
                if (steps > 100000000) {
                    std::printf("n=%u: a search does not end\n", n);
                    return 1;
                }
                uint64_t b = 0;
                const int s = ts.pass(s4, s1, &b);
                ++passes;
                if (s == 1) break;
                if (s == 2) {
                    for (uint32_t q = it.row; q <= last; ++q) col[q] = (uint32_t)((b >> (5u * (q - it.row))) & 31u);
                    uint64_t lo, hi;
                    pack(col, n, lo, hi);
                    if (!found.insert({lo, hi}).second) {
                        std::printf("n=%u: a board found twice\n", n);
                        return 1;
                    }
                    const uint32_t w = sym_check_packed(c, lo, hi);
                    got_classes += w != 0u;
                    got_weight += w;
                }
            }
            if (found != g.second) {
                std::printf("n=%u: record found %zu boards, %zu wrote it\n", n, found.size(), g.second.size());
                return 1;
            }
        }
    }
    std::printf("n=%u items=%zu records=%llu tie boards=%llu passes=%llu classes=%llu weight=%llu\n", n,
                items.size(), (unsigned long long)records, (unsigned long long)boards, (unsigned long long)passes,
                (unsigned long long)got_classes, (unsigned long long)got_weight);
    if (got_classes != want_classes || got_weight != want_weight || records == 0) {
        std::printf("n=%u: classes %llu / %llu, weight %llu / %llu\n", n, (unsigned long long)got_classes,
                    (unsigned long long)want_classes, (unsigned long long)got_weight, (unsigned long long)want_weight);
        return 1;
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const uint32_t lo = argc > 1 ? (uint32_t)std::atoi(argv[1]) : 10u, hi = argc > 2 ? (uint32_t)std::atoi(argv[2]) : 13u;
    for (uint32_t n = lo; n <= hi; ++n)
        if (check(n)) return 1;
    std::printf("ok\n");
    return 0;
}
