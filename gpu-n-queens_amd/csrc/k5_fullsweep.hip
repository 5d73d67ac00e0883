// k5_fullsweep.hip -- full DoubleSweep counting (K5): propagation over every
// open row and column, branching centre-out.
//
// The reference searches with unit propagation of the current row only
// (DoubleSweep-light, GPU_NQueens.cu:392-408) and names the full algorithm as
// future work (README.md:6, :22, :40-41); its unbuilt attempt at it walks the
// rows centre-out (CALC_CURRENT_ROW, GPU_NQueens.cu:207; :214-336). The search
// counted here:
//   propagate: repeat { every open row, ascending (the open rows when the pass
//     starts): no free cell kills the board, a single free cell gets its
//     queen; then the same over the open columns (a column's free cells are
//     its cells in open rows) } until a row pass and the column pass after it
//     place nothing;
//   branch on the open row nearest m = (n-1)/2, in the order m, m+1, m-1,
//     m+2, ..., its free columns ascending;
//   the root branches on row m: left-half columns weigh 2, the middle column
//     of odd n weighs 1 (mirror halving as in GPU_NQueens.cu:354-357).
// Every count is the reference's; nodes (the placements executed: children and
// propagated queens) are this search's own and differ from the light search's.
//
// Board state in registers: P (rows placed), C (columns placed), D1 (bit r+c)
// and D2 (bit c-r+n-1) as 64-bit masks (2n-1 <= 41 bits for n <= 21). Free
// cells of row r: ~(C | D1 >> r | D2 >> (n-1-r)); of column c, as a row
// mask: ~(P | D1 >> c | rev_n(D2 >> c)) (rev_n reverses the low n bits).
//
// Pipeline: the host places the root's children; fs_expand_kernel levels (one
// thread per state: propagate, then write the branch row's children through a
// block-aggregated cursor) run until a level holds enough states; the
// persistent fs_drain_kernel then runs one search per lane with a stack of
// branch points in LDS, lanes taking states per wave from a queue head.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "nqm_internal.h"

namespace nqm {
namespace {

constexpr uint32_t kFsMaxN = 21;         // D1 / D2 in 41 bits; P and C packed beside them in LDS
constexpr uint32_t kFsPackShift = 9;     // bits of D1 / D2 above bit 31 (2n-1-32 <= 9)
constexpr uint32_t kFsExpandThreads = 256;
constexpr uint32_t kFsDrainThreads = 64; // one wave per workgroup
constexpr uint32_t kFsAccSlots = 64;     // expansion sums spread over 64 counters (256 B apart)
constexpr uint32_t kFsMaxLevels = 16;

struct FsItem {
    uint32_t P, C, d1lo, d1hi, d2lo, d2hi, w, pad;
};

struct FsState {
    uint32_t P, C;
    uint64_t D1, D2;
};

__host__ __device__ __forceinline__ uint32_t fs_ctz(uint32_t v) { return (uint32_t)__builtin_ctz(v); }

__host__ __device__ __forceinline__ uint32_t fs_row_free(const FsState& s, uint32_t r, uint32_t n, uint32_t M) {
    return ~(s.C | (uint32_t)(s.D1 >> r) | (uint32_t)(s.D2 >> (n - 1u - r))) & M;
}

__host__ __device__ __forceinline__ uint32_t fs_col_free(const FsState& s, uint32_t c, uint32_t n, uint32_t M) {
    // D2 bit c+n-1-r blocks row r: bit n-1-r of the window D2 >> c, reversed
    const uint32_t rev = __builtin_bitreverse32((uint32_t)(s.D2 >> c)) >> (32u - n);
    return ~(s.P | (uint32_t)(s.D1 >> c) | rev) & M;
}

__host__ __device__ __forceinline__ void fs_place(FsState& s, uint32_t r, uint32_t c, uint32_t n) {
    s.P |= 1u << r;
    s.C |= 1u << c;
    s.D1 |= 1ull << (r + c);
    s.D2 |= 1ull << (c + n - 1u - r);
}

// Returns false when a line has no free cell; the placements it makes are
// added to nodes.
__host__ __device__ __forceinline__ bool fs_propagate(FsState& s, uint32_t n, uint32_t M, uint32_t& nodes) {
    for (;;) {
        bool placed = false;
        for (uint32_t o = ~s.P & M; o; o &= o - 1u) {
            const uint32_t r = fs_ctz(o);
            const uint32_t f = fs_row_free(s, r, n, M);
            if (!f) return false;
            if (!(f & (f - 1u))) {
                fs_place(s, r, fs_ctz(f), n);
                ++nodes;
                placed = true;
            }
        }
        for (uint32_t o = ~s.C & M; o; o &= o - 1u) {
            const uint32_t c = fs_ctz(o);
            const uint32_t g = fs_col_free(s, c, n, M);
            if (!g) return false;
            if (!(g & (g - 1u))) {
                fs_place(s, fs_ctz(g), c, n);
                ++nodes;
                placed = true;
            }
        }
        if (!placed) return true;
    }
}

// The open row nearest m = (n-1)/2, the row below first on a tie.
__host__ __device__ __forceinline__ uint32_t fs_branch_row(uint32_t P, uint32_t n, uint32_t M) {
    const uint32_t m = (n - 1u) >> 1;
    const uint32_t open = ~P & M;
    const uint32_t down = open >> m, up = open & ((1u << m) - 1u);
    const uint32_t dd = down ? fs_ctz(down) : 64u;
    const uint32_t du = up ? m - (31u - (uint32_t)__builtin_clz(up)) : 64u;
    return dd <= du ? m + dd : m - du;
}

__host__ __device__ __forceinline__ FsState fs_load(const FsItem& it) {
    FsState s;
    s.P = it.P;
    s.C = it.C;
    s.D1 = ((uint64_t)it.d1hi << 32) | it.d1lo;
    s.D2 = ((uint64_t)it.d2hi << 32) | it.d2lo;
    return s;
}

__host__ __device__ __forceinline__ FsItem fs_item(const FsState& s, uint32_t w) {
    FsItem it;
    it.P = s.P;
    it.C = s.C;
    it.d1lo = (uint32_t)s.D1;
    it.d1hi = (uint32_t)(s.D1 >> 32);
    it.d2lo = (uint32_t)s.D2;
    it.d2hi = (uint32_t)(s.D2 >> 32);
    it.w = w;
    it.pad = 0;
    return it;
}

// One level: each thread propagates one state; a complete board adds its
// weight, a stable one writes its branch row's children (each a placement)
// at the block's range of the cursor. Sums go to acc[slot * 32 + {0, 1}]
// (solutions, placements). A level larger than cap sets err bit 1 and
// writes nothing past cap.
__global__ void __launch_bounds__(kFsExpandThreads)
fs_expand_kernel(const FsItem* __restrict__ in, uint32_t count, uint32_t n, FsItem* __restrict__ out, uint32_t cap,
                 uint32_t* __restrict__ cursor, unsigned long long* __restrict__ acc, uint32_t* __restrict__ err) {
    __shared__ uint32_t wave_tot[kFsExpandThreads / 64];
    __shared__ uint32_t block_base;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
    const uint32_t i = blockIdx.x * kFsExpandThreads + tid;
    const uint32_t M = (uint32_t)((1ull << n) - 1ull);
    FsState s{};
    uint32_t w = 0, r = 0, f = 0, nodes = 0;
    uint64_t sols = 0;
    if (i < count) {
        const FsItem it = in[i];
        s = fs_load(it);
        w = it.w;
        if (fs_propagate(s, n, M, nodes)) {
            if (s.P == M) {
                sols = w;
            } else {
                r = fs_branch_row(s.P, n, M);
                f = fs_row_free(s, r, n, M);
            }
        }
    }
    const uint32_t k = (uint32_t)__popc(f);
    nodes += k;
    // block exclusive scan of k
    uint32_t inc = k;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(inc, o, 64);
        if ((int)lane >= o) inc += v;
    }
    if (lane == 63u) wave_tot[wid] = inc;
    __syncthreads();
    if (tid == 0) {
        uint32_t t = 0;
        for (uint32_t j = 0; j < kFsExpandThreads / 64; ++j) {
            const uint32_t v = wave_tot[j];
            wave_tot[j] = t;
            t += v;
        }
        block_base = t ? atomicAdd(cursor, t) : 0u;
    }
    __syncthreads();
    uint32_t off = block_base + wave_tot[wid] + inc - k;
    if (k) {
        if ((uint64_t)off + k > cap) {
            atomicOr(err, 1u);
        } else {
            for (uint32_t g = f; g; g &= g - 1u) {
                FsState t = s;
                fs_place(t, r, fs_ctz(g), n);
                out[off++] = fs_item(t, w);
            }
        }
    }
    uint64_t nd = nodes;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        sols += __shfl_xor(sols, o, 64);
        nd += __shfl_xor(nd, o, 64);
    }
    if (lane == 0 && (sols | nd)) {
        const uint32_t slot = (blockIdx.x * (kFsExpandThreads / 64) + wid) % kFsAccSlots;
        atomicAdd(&acc[slot * 32u], (unsigned long long)sols);
        atomicAdd(&acc[slot * 32u + 1u], (unsigned long long)nd);
    }
}

struct FsDrainArgs {
    const FsItem* items;
    uint32_t count;
    uint32_t n;
    uint32_t stack;         // branch points a lane can hold
    uint32_t refill_idle;   // a wave takes states when this many lanes are idle
    uint32_t* head;         // states handed out: kHeadSegs segments of the queue, a head each kHeadStride apart
    uint64_t* wave_part;    // per workgroup {solutions, placements}
    uint32_t* err;          // bit 2: a stack overflowed
};

__device__ __forceinline__ uint32_t fs_rank(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// ((hi:lo) >> s) & 0xffffffff for s < 32 (one v_alignbit on the device).
__host__ __device__ __forceinline__ uint32_t fs_align(uint32_t hi, uint32_t lo, uint32_t s) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbit(hi, lo, s);
#else
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> s);
#endif
}

// One lane of the drain: the state of its search, advanced one step at a time
// so that every lane of a wave does the same work per step (up to
// kFsLinesPerStep lines evaluated, at most one pass change, branch or
// backtrack).
struct FsLane {
    uint32_t P, C, d1lo, d1hi, d2lo, d2hi;
    uint32_t o;        // lines of the current pass not yet evaluated
    uint32_t col;      // 1: the pass is over columns
    uint32_t changed;  // the row or column pass placed a queen
    uint32_t sp, w;
    bool act;
    uint64_t sols, nodes;
};

__host__ __device__ __forceinline__ void fs_lane_place(FsLane& L, uint32_t r, uint32_t c, uint32_t n) {
    L.P |= 1u << r;
    L.C |= 1u << c;
    const uint64_t d1 = 1ull << (r + c), d2 = 1ull << (c + n - 1u - r);
    L.d1lo |= (uint32_t)d1;
    L.d1hi |= (uint32_t)(d1 >> 32);
    L.d2lo |= (uint32_t)d2;
    L.d2hi |= (uint32_t)(d2 >> 32);
}

__host__ __device__ __forceinline__ void fs_lane_start(FsLane& L, const FsItem& it, uint32_t M) {
    L.P = it.P;
    L.C = it.C;
    L.d1lo = it.d1lo;
    L.d1hi = it.d1hi;
    L.d2lo = it.d2lo;
    L.d2hi = it.d2hi;
    L.w = it.w;
    L.o = ~it.P & M;
    L.col = 0;
    L.changed = 0;
    L.sp = 0;
    L.act = true;
}

// One step of an active lane. The stack of branch points: entry j at
// stk[j * stride] (D1 lo, D1 hi | P << 9, D2 lo, D2 hi | C << 9) and its
// untried columns at stkf[j * stride]. Returns false when the stack would
// overflow (the lane stops; the count is void).
#ifndef NQM_FS_LINES
#define NQM_FS_LINES 8
#endif
constexpr uint32_t kFsLinesPerStep = NQM_FS_LINES;

__host__ __device__ __forceinline__ bool fs_lane_step(FsLane& L, uint32_t n, uint32_t M, uint32_t stack, uint4* stk,
                                                      uint32_t* stkf, uint32_t stride) {
    // up to kFsLinesPerStep lines of the current pass, in order, stopping at
    // a dead line or the pass's end: the pass change, the branch and the pop
    // below then run once per several line evaluations, not once per line
    bool dead = false;
#pragma unroll
    for (uint32_t i = 0; i < kFsLinesPerStep; ++i) {
        if (L.o && !dead) {
            const uint32_t line = fs_ctz(L.o);
            L.o &= L.o - 1u;
            const uint32_t a = fs_align(L.d1hi, L.d1lo, line);
            const uint32_t x = L.col ? L.P : L.C;
            const uint32_t b = L.col ? __builtin_bitreverse32(fs_align(L.d2hi, L.d2lo, line)) >> (32u - n)
                                     : fs_align(L.d2hi, L.d2lo, n - 1u - line);
            const uint32_t F = ~(x | a | b) & M;
            if (F == 0u) {
                dead = true;
            } else if ((F & (F - 1u)) == 0u) {
                const uint32_t q = fs_ctz(F);
                fs_lane_place(L, L.col ? q : line, L.col ? line : q, n);
                ++L.nodes;
                L.changed = 1u;
            }
        }
    }
    bool child = false, popped = false;
    uint32_t r = 0, f = 0, j = 0;
    if (!dead && L.o == 0u) {
        if (!L.col) {
            L.col = 1u;
            L.o = ~L.C & M;
        } else if (L.changed) {
            L.col = 0u;
            L.changed = 0u;
            L.o = ~L.P & M;
        } else if (L.P == M) {
            L.sols += L.w;
            dead = true;
        } else {
            r = fs_branch_row(L.P, n, M);
            f = ~(L.C | fs_align(L.d1hi, L.d1lo, r) | fs_align(L.d2hi, L.d2lo, n - 1u - r)) & M;
            child = true;
        }
    }
    if (dead) {
        if (L.sp == 0u) {
            L.act = false;
        } else {
            j = (L.sp - 1u) * stride;
            const uint4 e = stk[j];
            f = stkf[j];
            const uint32_t pm = (1u << kFsPackShift) - 1u;
            L.d1lo = e.x;
            L.d1hi = e.y & pm;
            L.P = e.y >> kFsPackShift;
            L.d2lo = e.z;
            L.d2hi = e.w & pm;
            L.C = e.w >> kFsPackShift;
            r = fs_branch_row(L.P, n, M);
            child = true;
            popped = true;
        }
    }
    if (child) {
        const uint32_t bit = f & (0u - f);
        f ^= bit;
        if (popped) {
            if (f)
                stkf[j] = f;
            else
                --L.sp;
        } else if (f) {  // always: a stable open row has two free cells at least
            if (L.sp >= stack) {
                L.act = false;
                return false;
            }
            j = L.sp * stride;
            stk[j] = make_uint4(L.d1lo, L.d1hi | (L.P << kFsPackShift), L.d2lo, L.d2hi | (L.C << kFsPackShift));
            stkf[j] = f;
            ++L.sp;
        }
        fs_lane_place(L, r, fs_ctz(bit), n);
        ++L.nodes;
        L.col = 0u;
        L.changed = 0u;
        L.o = ~L.P & M;
    }
    return true;
}

// LDS: entry j of lane t at uint4 [j * 64 + t], its untried columns at
// u32 [stack * 64 * 4 + j * 64 + t].
__global__ void __launch_bounds__(kFsDrainThreads) fs_drain_kernel(FsDrainArgs a) {
    extern __shared__ uint4 fs_lds[];
    const uint32_t lane = threadIdx.x;
    uint4* stk = fs_lds + lane;
    uint32_t* stkf = reinterpret_cast<uint32_t*>(fs_lds + a.stack * kFsDrainThreads) + lane;
    const uint32_t n = a.n, M = (uint32_t)((1ull << n) - 1ull);
    FsLane L{};
    bool drained = false;
    // the queue in kHeadSegs segments with a head each (no head that every
    // wave of the chip updates); a wave starts at its own and moves on when it
    // runs out
    auto seg_lo = [&](uint32_t sg) { return (uint32_t)((uint64_t)a.count * sg / kHeadSegs); };
    uint32_t seg = blockIdx.x % kHeadSegs, seg_done = 0;
    for (;;) {
        const uint64_t idle = __ballot(!L.act);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (!drained && (nidle >= a.refill_idle || nidle == 64u)) {
            uint32_t base = 0, got = 0;
            while (got == 0 && !drained) {  // wave-uniform, at most kHeadSegs rounds
                const uint32_t lo = seg_lo(seg), len = seg_lo(seg + 1u) - lo;
                uint32_t g = 0;
                if (lane == 0) g = atomicAdd(a.head + seg * kHeadStride, nidle);
                g = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl((int)g, 0, 64));
                if (g < len) {
                    base = lo + g;
                    got = min(nidle, len - g);
                }
                if (got < nidle) {  // this segment is used up
                    if (++seg_done == kHeadSegs) drained = true;
                    seg = seg + 1u == kHeadSegs ? 0u : seg + 1u;
                }
            }
            if (!L.act) {
                const uint32_t r = fs_rank(idle);
                if (r < got) fs_lane_start(L, a.items[base + r], M);
            }
        }
        if (drained && __ballot(L.act) == 0) break;
        if (L.act && !fs_lane_step(L, n, M, a.stack, stk, stkf, kFsDrainThreads)) atomicOr(a.err, 2u);
    }
    uint64_t sols = L.sols, nodes = L.nodes;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        sols += __shfl_xor(sols, o, 64);
        nodes += __shfl_xor(nodes, o, 64);
    }
    if (lane == 0) {
        a.wave_part[2u * blockIdx.x] = sols;
        a.wave_part[2u * blockIdx.x + 1u] = nodes;
    }
}

size_t fs_lds_bytes(uint32_t stack) { return (size_t)stack * kFsDrainThreads * 20u; }

}  // namespace
}  // namespace nqm

using namespace nqm;

extern "C" int nqm_count_fullsweep_ex(nqm_ctx* ctx, uint32_t n, uint32_t target_states, uint64_t* total,
                                      nqm_stats* stats) {
    if (!ctx) return NQM_EINVAL;
    if (n < 1 || n > kMaxN) return fail(ctx, NQM_EINVAL, "n = %u outside 1..%u", n, kMaxN);
    if (n > kFsMaxN) return fail(ctx, NQM_EUNSUPPORTED, "full sweep counts n <= %u (n = %u)", kFsMaxN, n);
    const double t0 = now_ms();
    NQM_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    // Automatic (target_states 0): expand until a level holds 2^19 states
    // and a lane's stack needs at most kFsRowsGoal rows, while levels stay
    // under 2^26 states. Measured on the drain at N=15..18: more stack rows
    // cost resident waves (N=18: 79.7 ms at 10 rows, 69.4 at 9), more levels
    // past that cost a state's start for too few placements (N=16: 2.3 ms
    // at 8 rows, 5.3 at 7).
    constexpr uint32_t kFsRowsGoal = 9, kFsMinStates = 1u << 19, kFsMaxStates = 1u << 26;
    const bool automatic = target_states == 0;
    const uint32_t target = automatic ? kFsMinStates : target_states;
    FsWork& W = ctx->fs;

    // the root's children: row m, left-half columns weigh 2, the middle of odd n 1
    std::vector<FsItem> roots;
    const uint32_t m = (n - 1u) / 2u;
    for (uint32_t c = 0; c < n / 2u + (n & 1u); ++c) {
        FsState s{};
        fs_place(s, m, c, n);
        roots.push_back(fs_item(s, c < n / 2u ? 2u : 1u));
    }
    uint64_t host_nodes = roots.size();

    int rc;
    // counters: acc (kFsAccSlots x 256 B), cursor at byte 16384, err at 16388, the queue's
    // heads from byte 16640 (kHeadSegs x kHeadStride u32)
    const size_t ctr_bytes = 16640 + (size_t)kHeadSegs * kHeadStride * 4u;
    if ((rc = ensure_dev(ctx, W.ctr, ctr_bytes)) != NQM_OK) return rc;
    unsigned long long* d_acc = static_cast<unsigned long long*>(W.ctr.ptr);
    uint32_t* d_cursor = reinterpret_cast<uint32_t*>(static_cast<char*>(W.ctr.ptr) + 16384);
    uint32_t* d_err = d_cursor + 1;
    uint32_t* d_head = reinterpret_cast<uint32_t*>(static_cast<char*>(W.ctr.ptr) + 16640);
    NQM_HIP(ctx, hipMemsetAsync(W.ctr.ptr, 0, ctr_bytes, st));
    if ((rc = ensure_dev(ctx, W.lvl[0], roots.size() * sizeof(FsItem))) != NQM_OK) return rc;
    NQM_HIP(ctx, hipMemcpyAsync(W.lvl[0].ptr, roots.data(), roots.size() * sizeof(FsItem), hipMemcpyHostToDevice, st));

    NQM_HIP(ctx, hipEventRecord(ctx->ev_x0, st));
    uint32_t cur = 0, count = (uint32_t)roots.size(), levels = 0;
    // a state at level L holds at least L + 1 queens; a lane's stack holds
    // one branch point per open row but the last
    auto stack_rows_after = [&](uint32_t lv) { return lv + 2u < n ? n - lv - 2u : 1u; };  // placed_min = lv + 1
    while (count > 0 && levels < kFsMaxLevels &&
           (count < target || (automatic && stack_rows_after(levels) > kFsRowsGoal && count < kFsMaxStates))) {
        const uint64_t cap64 = (uint64_t)count * n;
        if (cap64 > 0xffffffffull) return fail(ctx, NQM_ERANGE, "level of %u states too large", count);
        const uint32_t cap = (uint32_t)cap64;
        if ((rc = ensure_dev(ctx, W.lvl[cur ^ 1u], (size_t)cap * sizeof(FsItem))) != NQM_OK) return rc;
        NQM_HIP(ctx, hipMemsetAsync(d_cursor, 0, 4, st));
        const uint32_t grid = (count + kFsExpandThreads - 1u) / kFsExpandThreads;
        hipLaunchKernelGGL(fs_expand_kernel, dim3(grid), dim3(kFsExpandThreads), 0, st,
                           static_cast<const FsItem*>(W.lvl[cur].ptr), count, n,
                           static_cast<FsItem*>(W.lvl[cur ^ 1u].ptr), cap, d_cursor, d_acc, d_err);
        NQM_HIP(ctx, hipGetLastError());
        uint32_t next = 0, err = 0;
        NQM_HIP(ctx, hipMemcpyAsync(&next, d_cursor, 4, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipMemcpyAsync(&err, d_err, 4, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipStreamSynchronize(st));
        if (err) return fail(ctx, NQM_EDEVICE, "expansion level %u overflowed (err %u)", levels, err);
        cur ^= 1u;
        count = next;
        ++levels;
    }
    const uint32_t placed_min = levels + 1u;
    const uint32_t stack = placed_min + 1u < n ? n - placed_min - 1u : 1u;

    // drain: one wave per workgroup, as many as LDS and registers allow
    const size_t lds = fs_lds_bytes(stack);
    int per_cu = 0;
    NQM_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fs_drain_kernel, kFsDrainThreads, lds));
    if (per_cu < 1) return fail(ctx, NQM_ERANGE, "drain stack of %u rows does not fit LDS", stack);
    const uint32_t blocks = count ? (uint32_t)std::min<uint64_t>((uint64_t)per_cu * ctx->prop.multiProcessorCount,
                                                                (count + kFsDrainThreads - 1u) / kFsDrainThreads)
                                  : 0u;
    if ((rc = ensure_dev(ctx, W.wave, std::max<size_t>(blocks, 1u) * 16u)) != NQM_OK) return rc;
    float search_ms = 0.f, all_ms = 0.f;
    NQM_HIP(ctx, hipEventRecord(ctx->ev_start, st));
    if (blocks) {
        FsDrainArgs a;
        a.items = static_cast<const FsItem*>(W.lvl[cur].ptr);
        a.count = count;
        a.n = n;
        a.stack = stack;
        a.refill_idle = 4u;  // few lanes idle at a time (the queue's heads are segmented)
        a.head = d_head;
        a.wave_part = static_cast<uint64_t*>(W.wave.ptr);
        a.err = d_err;
        hipLaunchKernelGGL(fs_drain_kernel, dim3(blocks), dim3(kFsDrainThreads), lds, st, a);
        NQM_HIP(ctx, hipGetLastError());
    }
    NQM_HIP(ctx, hipEventRecord(ctx->ev_stop, st));
    NQM_HIP(ctx, hipEventRecord(ctx->ev_x1, st));
    std::vector<uint64_t> wp((size_t)blocks * 2u);
    std::vector<unsigned long long> acc((size_t)kFsAccSlots * 32u);
    uint32_t err = 0;
    if (blocks) NQM_HIP(ctx, hipMemcpyAsync(wp.data(), W.wave.ptr, wp.size() * 8u, hipMemcpyDeviceToHost, st));
    NQM_HIP(ctx, hipMemcpyAsync(acc.data(), d_acc, acc.size() * 8u, hipMemcpyDeviceToHost, st));
    NQM_HIP(ctx, hipMemcpyAsync(&err, d_err, 4, hipMemcpyDeviceToHost, st));
    NQM_HIP(ctx, hipStreamSynchronize(st));
    if (err & 2u) return fail(ctx, NQM_EDEVICE, "a drain stack overflowed (%u rows)", stack);
    if (err) return fail(ctx, NQM_EDEVICE, "full sweep failed (err %u)", err);
    NQM_HIP(ctx, hipEventElapsedTime(&search_ms, ctx->ev_start, ctx->ev_stop));
    NQM_HIP(ctx, hipEventElapsedTime(&all_ms, ctx->ev_x0, ctx->ev_x1));
    uint64_t sols = 0, nodes = host_nodes;
    for (uint32_t s = 0; s < kFsAccSlots; ++s) {
        sols += acc[(size_t)s * 32u];
        nodes += acc[(size_t)s * 32u + 1u];
    }
    for (uint32_t b = 0; b < blocks; ++b) {
        sols += wp[2u * b];
        nodes += wp[2u * b + 1u];
    }
    if (total) *total = sols;
    if (stats) {
        *stats = nqm_stats{};
        stats->total = sols;
        stats->nodes = nodes;
        stats->kernel_ms = all_ms;
        stats->search_ms = search_ms;
        stats->wall_ms = now_ms() - t0;
        stats->launches = levels + (blocks ? 1u : 0u);
        stats->blocks = blocks;
        stats->rounds = blocks ? 1u : 0u;
        stats->lanes = (uint64_t)blocks * kFsDrainThreads;
    }
    return NQM_OK;
}

extern "C" int nqm_count_fullsweep(nqm_ctx* ctx, uint32_t n, uint64_t* total, nqm_stats* stats) {
    return nqm_count_fullsweep_ex(ctx, n, 0u, total, stats);
}
