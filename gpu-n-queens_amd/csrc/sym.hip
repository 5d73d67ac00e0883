// sym.hip -- symmetry-reduced counting ("transformation elimination"), one
// thread per item (K3).
//
// The reference lists it as future work (README.md:38-45) and counts every
// solution of the mirror-halved tree. This path counts each class of
// solutions under the 8 symmetries of the board once, by the corner /
// non-corner method of sym_rules.h, and weights it by its orbit size (2, 4
// or 8). The host enumerates the first rows (same rules) into items; each
// lane of k3_sym_kernel searches one item with an LDS stack and keeps its
// board in LDS for the rotation check. Totals equal the reference's counts;
// the node numbers are the reduced search's own. The persistent drain of
// k4_sym.hip counts the same tree faster; this kernel stays as its
// reference on the device and serves the boards the drain does not take.
#include <algorithm>
#include <new>
#include <vector>

#include "nqm_internal.h"
#include "sym_rules.h"

namespace nqm {
namespace {

// One item: the case, the next row y and its masks, the columns of rows < y.
struct SymItem {
    uint32_t l, d, r;
    uint8_t y, b1, corner, n;
    uint8_t col[32];
};
static_assert(sizeof(SymItem) == 48, "item layout");

constexpr uint32_t kSymLanes = 64;  // one wave per workgroup
constexpr size_t kSymMaxItems = (size_t)1 << 26;  // 3 GiB of host items; N=18 makes 0.9 M

// LDS per workgroup: the stack, [row][lane] x 16 B {f, l, d, r} for rows
// y0..last-1, then the board, [row][lane] x 1 B (column) for rows 0..31.
size_t sym_lds_bytes(uint32_t rows) { return (size_t)rows * kSymLanes * 16u + 32u * kSymLanes; }

struct LdsBoard {
    uint8_t* brd;
    uint32_t lane;
    __device__ uint32_t operator()(int row) const { return 1u << brd[(uint32_t)row * kSymLanes + lane]; }
    __device__ void set(uint32_t row, uint32_t col) { brd[row * kSymLanes + lane] = (uint8_t)col; }
};

struct SymCounts {
    uint64_t c2 = 0, c4 = 0, c8 = 0, nodes = 0;
};

// The search of one item: rows y0..last, an explicit stack of the rows above
// the current one (stack(k) is row y0 + k, k < last - y0) and the board
// (board.set / board(row), rows 0..last). Shared by the kernel (LDS) and
// host checks of the same loop.
template <class Stack, class Board>
__host__ __device__ inline void sym_search(const SymItem& it, Stack& stack, Board& board, SymCounts& k) {
    const SymCase c = sym_case(it.n, it.b1, it.corner);
    const uint32_t y0 = it.y;
    for (uint32_t row = 0; row < y0; ++row) board.set(row, it.col[row]);
    uint32_t y = y0, l = it.l, d = it.d, r = it.r;
    uint32_t f = sym_free(c, y, l, d, r);
    uint32_t placed = 0;  // an item's subtree is at most (n - y0)! placements: 11 rows left fit
    for (;;) {
        if (f == 0u) {  // row exhausted: back to the row above (within the item)
            if (y == y0) break;
            --y;
            const uint4 e = stack.get(y - y0);
            f = e.x;
            l = e.y;
            d = e.z;
            r = e.w;
            continue;
        }
        const uint32_t bit = f & (0u - f);
        f ^= bit;
        ++placed;
        board.set(y, (uint32_t)__builtin_ctz(bit));
        if (y == c.last) {  // a full board (the last row has one free column at most)
            if (c.corner) {
                ++k.c8;
            } else {
                const uint32_t m = sym_check(c, board);
                k.c2 += m == 2u;
                k.c4 += m == 4u;
                k.c8 += m == 8u;
            }
            continue;
        }
        stack.put(y - y0, make_uint4(f, l, d, r));
        l = (l | bit) << 1;
        d |= bit;
        r = (r | bit) >> 1;
        ++y;
        f = sym_free(c, y, l, d, r);
    }
    k.nodes += placed;
}

struct LdsStack {
    uint4* stk;
    uint32_t lane;
    __device__ uint4 get(uint32_t k) const { return stk[k * kSymLanes + lane]; }
    __device__ void put(uint32_t k, uint4 e) { stk[k * kSymLanes + lane] = e; }
};

__global__ void __launch_bounds__(kSymLanes)
k3_sym_kernel(const SymItem* __restrict__ items, uint32_t count, uint32_t rows, uint64_t* __restrict__ part) {
    extern __shared__ uint4 lds_sym[];
    const uint32_t lane = threadIdx.x;
    const uint32_t i = blockIdx.x * kSymLanes + lane;
    SymCounts k;
    if (i < count) {
        LdsStack stack{lds_sym, lane};
        LdsBoard board{reinterpret_cast<uint8_t*>(lds_sym + (size_t)rows * kSymLanes), lane};
        sym_search(items[i], stack, board, k);
    }
    uint64_t c2 = k.c2, c4 = k.c4, c8 = k.c8, nodes = k.nodes;
    for (int o = 32; o > 0; o >>= 1) {
        c2 += __shfl_xor(c2, o, 64);
        c4 += __shfl_xor(c4, o, 64);
        c8 += __shfl_xor(c8, o, 64);
        nodes += __shfl_xor(nodes, o, 64);
    }
    if (lane == 0) {
        uint64_t* const p = part + (size_t)blockIdx.x * 4u;
        p[0] = c2;
        p[1] = c4;
        p[2] = c8;
        p[3] = nodes;
    }
}

// Host enumeration of the first rows; rows that reach the last row on the
// host (small boards) are counted here.
struct SymHost {
    uint32_t depth;
    size_t max_items;       // the enumeration stops (overflow) past this many items
    bool overflow = false;
    std::vector<SymItem> items;
    uint64_t c2 = 0, c4 = 0, c8 = 0, nodes = 0;
    uint8_t col[32];
};

struct HostBoard {
    const uint8_t* col;
    uint32_t operator()(int row) const { return 1u << col[row]; }
};

void sym_host_rec(SymHost& h, const SymCase& c, uint32_t y, uint32_t l, uint32_t d, uint32_t r) {
    if (h.overflow) return;
    if (y == h.depth && y < c.last) {
        if (h.items.size() >= h.max_items) {
            h.overflow = true;
            return;
        }
        SymItem it{};
        it.l = l;
        it.d = d;
        it.r = r;
        it.y = (uint8_t)y;
        it.b1 = (uint8_t)c.b1;
        it.corner = (uint8_t)c.corner;
        it.n = (uint8_t)c.n;
        std::copy(h.col, h.col + y, it.col);
        h.items.push_back(it);
        return;
    }
    uint32_t f = sym_free(c, y, l, d, r);
    while (f) {
        const uint32_t bit = f & (0u - f);
        f ^= bit;
        ++h.nodes;
        h.col[y] = (uint8_t)__builtin_ctz(bit);
        if (y == c.last) {
            if (c.corner) {
                ++h.c8;
            } else {
                const uint32_t k = sym_check(c, HostBoard{h.col});
                h.c2 += k == 2u;
                h.c4 += k == 4u;
                h.c8 += k == 8u;
            }
            continue;
        }
        sym_host_rec(h, c, y + 1, (l | bit) << 1, d | bit, (r | bit) >> 1);
    }
}

}  // namespace
}  // namespace nqm

using namespace nqm;

namespace {

// K3: host items, one thread each.
int k3_count(nqm_ctx* ctx, uint32_t n, uint64_t* unique, uint64_t* total, nqm_stats* stats) {
    const double t0 = now_ms();
    uint64_t host_unique = 0, host_total = 0;
    SymHost h;
    if (n < 4) {  // too small for the bound rows: 1 (one class of one) for n = 1, none for 2 and 3
        host_unique = host_total = n == 1 ? 1u : 0u;
    } else {
        // Rows placed on the host: the device searches the last 11 rows, 12
        // from N=18 on. Measured: one row less at N=16 left 106 workgroups
        // (92 ms against 19); at N=18 one row more made 4.4 M items whose
        // enumeration and upload took 155 of 278 ms (158 ms this way).
        h.depth = n <= 11 ? n : std::min(n - 3u, std::max(2u, n < 18 ? n - 11u : n - 12u));
        // one thread per item, 64 per workgroup: the grid and the per-block
        // records stay within u32 and the host buffer within a few GiB
        h.max_items = kSymMaxItems;
        const uint32_t last = n - 1u;
        try {
            for (uint32_t b1 = 2; b1 < last; ++b1) {  // corner: rows 0 and 1 fixed
                const SymCase c = sym_case(n, b1, 1u);
                const uint32_t bit = 1u << b1;
                h.col[0] = 0;
                h.col[1] = (uint8_t)b1;
                h.nodes += 2;
                sym_host_rec(h, c, 2, (2u | bit) << 1, 1u | bit, bit >> 1);
            }
            for (uint32_t b1 = 1; b1 < last - b1; ++b1) {  // non-corner: row 0 fixed
                const SymCase c = sym_case(n, b1, 0u);
                const uint32_t bit = 1u << b1;
                h.col[0] = (uint8_t)b1;
                h.nodes += 1;
                sym_host_rec(h, c, 1, bit << 1, bit, bit >> 1);
            }
        } catch (const std::bad_alloc&) {
            return fail(ctx, NQM_ENOMEM, "host items of the symmetric count (n=%u)", n);
        }
        if (h.overflow)
            return fail(ctx, NQM_ERANGE, "n=%u: the symmetric count's host enumeration exceeds %zu items", n,
                        kSymMaxItems);
    }
    uint64_t c2 = h.c2, c4 = h.c4, c8 = h.c8, nodes = h.nodes;
    float ms = 0.f;
    const uint32_t count = (uint32_t)h.items.size();
    uint32_t blocks = 0;
    if (count) {
        hipStream_t st = ctx->stream;
        const uint32_t rows = n - 1u - h.depth;  // stack rows y0..last-1
        const size_t lds = sym_lds_bytes(rows);
        if (lds > ctx->lds_per_block) return fail(ctx, NQM_EUNSUPPORTED, "symmetric stack does not fit LDS");
        blocks = (count + kSymLanes - 1u) / kSymLanes;
        DevBuf d_items, d_part;
        struct Free {
            DevBuf& a;
            DevBuf& b;
            ~Free() {
                if (a.ptr) (void)hipFree(a.ptr);
                if (b.ptr) (void)hipFree(b.ptr);
            }
        } fr{d_items, d_part};
        int rc;
        if ((rc = ensure_dev(ctx, d_items, (size_t)count * sizeof(SymItem))) != NQM_OK) return rc;
        if ((rc = ensure_dev(ctx, d_part, (size_t)blocks * 32u)) != NQM_OK) return rc;
        NQM_HIP(ctx, hipMemcpyAsync(d_items.ptr, h.items.data(), (size_t)count * sizeof(SymItem),
                                    hipMemcpyHostToDevice, st));
        NQM_HIP(ctx, hipEventRecord(ctx->ev_start, st));
        hipLaunchKernelGGL(k3_sym_kernel, dim3(blocks), dim3(kSymLanes), lds, st,
                           static_cast<const SymItem*>(d_items.ptr), count, rows, static_cast<uint64_t*>(d_part.ptr));
        NQM_HIP(ctx, hipGetLastError());
        NQM_HIP(ctx, hipEventRecord(ctx->ev_stop, st));
        std::vector<uint64_t> part((size_t)blocks * 4u);
        NQM_HIP(ctx, hipMemcpyAsync(part.data(), d_part.ptr, part.size() * 8u, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipStreamSynchronize(st));
        NQM_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
        for (uint32_t b = 0; b < blocks; ++b) {
            c2 += part[b * 4u];
            c4 += part[b * 4u + 1u];
            c8 += part[b * 4u + 2u];
            nodes += part[b * 4u + 3u];
        }
    }
    const uint64_t u = host_unique + c2 + c4 + c8, t = host_total + 2 * c2 + 4 * c4 + 8 * c8;
    if (unique) *unique = u;
    if (total) *total = t;
    if (stats) {
        stats->total = t;
        stats->nodes = nodes;
        stats->kernel_ms = ms;
        stats->search_ms = ms;
        stats->wall_ms = now_ms() - t0;
        stats->launches = count ? 1u : 0u;
        stats->blocks = blocks;
    }
    return NQM_OK;
}

}  // namespace

extern "C" int nqm_count_symmetric_path(nqm_ctx* ctx, uint32_t n, uint32_t path, uint64_t* unique, uint64_t* total,
                                        nqm_stats* stats) {
    if (!ctx) return NQM_EINVAL;
    if (n < 1 || n > kMaxN) return fail(ctx, NQM_EINVAL, "n=%u outside 1..%u", n, kMaxN);
    if (path > NQM_SYM_DRAIN) return fail(ctx, NQM_EINVAL, "unknown symmetric path %u", path);
    if (path == NQM_SYM_DRAIN || (path == NQM_SYM_AUTO && k4_supports(n))) return k4_count(ctx, n, unique, total, stats);
    return k3_count(ctx, n, unique, total, stats);
}

extern "C" int nqm_count_symmetric(nqm_ctx* ctx, uint32_t n, uint64_t* unique, uint64_t* total, nqm_stats* stats) {
    return nqm_count_symmetric_path(ctx, n, NQM_SYM_AUTO, unique, total, stats);
}
