// nqm_items.h -- the work-item format shared by the queue kernels and host.
//
// A work item is one open search node: the queens already placed (as
// relative bitsets at row `row`) and the candidate columns of that row that
// are still to be tried. Every item is an independent subtree, so items can
// be split, reordered, spilled and resumed without changing any count.
#pragma once

#include <cstdint>

namespace nqm {

struct alignas(16) Item {
    uint32_t f;      // candidate columns still to try at `row` (never 0 in a queue)
    uint32_t cols;   // columns taken by rows 0..row-1
    uint32_t ld;     // columns attacked at `row` along down-left diagonals
    uint32_t rd;     // columns attacked at `row` along down-right diagonals
    uint32_t solve;  // index of the solve this item belongs to
    uint32_t meta;   // bits 0-4 n, bits 8-12 row, bit 16 weight x2 (mirror)
    uint32_t pad0, pad1;
};
static_assert(sizeof(Item) == 32, "Item is two 16-byte loads");

__host__ __device__ inline uint32_t item_meta(uint32_t n, uint32_t row, uint32_t wshift) {
    return (n & 31u) | ((row & 31u) << 8) | ((wshift & 1u) << 16);
}
__host__ __device__ inline uint32_t meta_n(uint32_t m) { return m & 31u; }
__host__ __device__ inline uint32_t meta_row(uint32_t m) { return (m >> 8) & 31u; }
__host__ __device__ inline uint32_t meta_wshift(uint32_t m) { return (m >> 16) & 1u; }

// Per-solve device accumulators (u64): weighted solutions and weighted
// (full-tree equivalent) nodes, followed by the executed node total.
struct SolveAcc {
    uint64_t* sols;   // [nsolves]
    uint64_t* nodes;  // [nsolves]
};

}  // namespace nqm
