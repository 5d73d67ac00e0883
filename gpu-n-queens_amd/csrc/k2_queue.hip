// k2_queue.hip -- K2: persistent work-queue counting for gfx950 (wave64).
//
// The reference statically gives each thread one pool state and lets a warp
// idle until its slowest lane finishes (GPU_NQueens.cu:338-360; the restart
// idea is future work in README.md:46-56). K2 keeps the same arithmetic (the
// per-lane placement step of K1, same counts, same node numbers) and changes
// only the scheduling:
//
//  * k2_expand: items at row r -> their children at row r+1, until each
//    solve's items reach its target row. One thread per item; children are
//    reserved with one atomic per wave (wave prefix sum). Dead children
//    (next row without a free cell) are dropped; every placement is counted
//    as a node of its solve. Item counts live in device memory, so a chain
//    of expand launches needs no host round trip (and can be graph-captured).
//
//  * k2_drain: a persistent grid. Lanes draw items from a global queue
//    (one atomicAdd per wave refill for all idle lanes, ranks by mbcnt).
//    When the queue is empty, idle lanes take work from busy lanes of the
//    same wave: a donor hands over the bottom entry of its LDS stack (the
//    shallowest open row, i.e. the largest untried subtree) and simply moves
//    its stack base up; the receiver reads that 16-byte entry straight from
//    the donor's LDS slot. Solves of different n mix freely: each lane holds
//    its own column mask. Counts are folded per lane and flushed to per-solve
//    u64 counters when the lane switches solve or exits.
//
//  * Optional pass budget: when a wave has run `budget` passes, each live
//    lane writes its open rows (current row + stack entries) to a spill
//    queue as independent items, and the wave exits. The host relaunches on
//    the spill queue (restart), which also lets new work join between rounds.

#include <algorithm>
#include <type_traits>

#include "nqm_internal.h"
#include "nqm_items.h"

namespace nqm {

constexpr uint32_t kK2Threads = 256;

__device__ __forceinline__ uint32_t lane_rank(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, uint32_t lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    return x;
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Add one u64 per lane into arr[solve] (lanes with add == 0 contribute
// nothing). When every contributing lane of the wave has the same solve --
// the common case, queues are grouped by solve -- the wave sums first and
// issues one atomic instead of 64 to the same address. All lanes must call.
__device__ __forceinline__ void add_per_solve(uint64_t* __restrict__ arr, uint32_t solve, uint64_t add,
                                              uint32_t lane) {
    const uint64_t act = __ballot(add != 0);
    if (act == 0) return;
    const uint32_t first = (uint32_t)__ffsll((long long)act) - 1u;
    const uint32_t s0 = __shfl(solve, first, 64);
    if (__ballot(add != 0 && solve != s0) == 0) {
        const uint64_t tot = wave_sum64(add);
        if (lane == first) atomicAdd((unsigned long long*)&arr[s0], (unsigned long long)tot);
    } else if (add) {
        atomicAdd((unsigned long long*)&arr[solve], (unsigned long long)add);
    }
}

// ---------------------------------------------------------------------------
// expand: one level of children for every item below its solve's target row
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k2_expand_kernel(const Item* __restrict__ in, const uint32_t* __restrict__ in_count,
                 Item* __restrict__ out, uint32_t* __restrict__ out_count, uint32_t out_cap,
                 const uint8_t* __restrict__ target_row, uint64_t* __restrict__ solve_nodes,
                 uint64_t* __restrict__ exec_nodes, uint32_t* __restrict__ err,
                 const float* __restrict__ pred, float thresh) {
    const uint32_t cnt = *in_count;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t step = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < cnt; base += step) {
        const uint32_t i = base + threadIdx.x;
        Item it{};
        uint32_t nch = 0, n = 0, row = 0, mask = 0, placed = 0;
        bool expand = false;
        if (i < cnt) {
            it = in[i];
            n = meta_n(it.meta);
            row = meta_row(it.meta);
            mask = (1u << n) - 1u;
            // by prediction (split the large subtrees, never the last rows)
            // or up to the solve's target row
            expand = pred ? (pred[i] > thresh && row + 2 < n) : (row < target_row[it.solve]);
            if (expand) {
                uint32_t f = it.f;
                placed = __popc(f);
                while (f) {
                    const uint32_t b = f & (0u - f);
                    f ^= b;
                    const uint32_t cf = ~((it.cols | b) | ((it.ld | b) << 1) | ((it.rd | b) >> 1)) & mask;
                    nch += (cf != 0u);
                }
            } else {
                nch = 1;
            }
        }
        // nodes placed by this level, weighted to full-tree numbers
        add_per_solve(solve_nodes, it.solve, expand ? (uint64_t)placed << meta_wshift(it.meta) : 0u, lane);
        const uint64_t wplaced = wave_sum64(placed);
        if (lane == 0 && wplaced) atomicAdd((unsigned long long*)exec_nodes, (unsigned long long)wplaced);
        const uint32_t incl = wave_incl_scan(nch, lane);
        const uint32_t total = __shfl(incl, 63, 64);
        uint32_t obase = 0;
        if (lane == 0 && total) obase = atomicAdd(out_count, total);
        obase = __shfl(obase, 0, 64);
        uint32_t o = obase + incl - nch;
        if (nch && (uint64_t)obase + total > out_cap) {
            atomicOr(err, 2u);
            continue;
        }
        if (i < cnt) {
            if (!expand) {
                out[o] = it;
            } else {
                uint32_t f = it.f;
                const uint32_t ws = meta_wshift(it.meta);
                while (f) {
                    const uint32_t b = f & (0u - f);
                    f ^= b;
                    Item c;
                    c.cols = it.cols | b;
                    c.ld = ((it.ld | b) << 1) & mask;
                    c.rd = (it.rd | b) >> 1;
                    c.f = ~(c.cols | c.ld | c.rd) & mask;
                    c.solve = it.solve;
                    c.meta = item_meta(n, row + 1, ws);
                    c.pad0 = c.pad1 = 0;
                    if (c.f) out[o++] = c;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// drain: persistent queue + intra-wave donation (+ optional spill budget)
// ---------------------------------------------------------------------------
struct DrainArgs {
    const Item* items;
    const uint32_t* item_count;
    uint32_t* head;
    Item* spill;
    uint32_t* spill_count;
    uint32_t spill_cap;
    uint64_t* solve_sols;
    uint64_t* solve_nodes;
    uint64_t* exec_nodes;
    uint32_t* err;
    uint32_t stack_rows;
    uint32_t refill_idle;
    uint32_t donate;
    uint32_t pad;
    uint64_t budget;  // 0: drain everything
};

__device__ __forceinline__ void flush_acc(const DrainArgs& a, uint32_t solve, uint64_t& sols,
                                          uint64_t& nodes) {
    if (sols) atomicAdd((unsigned long long*)&a.solve_sols[solve], (unsigned long long)sols);
    if (nodes) atomicAdd((unsigned long long*)&a.solve_nodes[solve], (unsigned long long)nodes);
    sols = 0;
    nodes = 0;
}

// kWide: u64 solution counter per lane, else u32 (allowed by the host only
// when every solve's board has fewer than 2^32 solutions, n <= 18).
template <bool kWide>
__global__ void __launch_bounds__(kK2Threads) k2_drain_kernel(DrainArgs a) {
    typedef typename std::conditional<kWide, uint64_t, uint32_t>::type sol_t;
    extern __shared__ uint4 lds[];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    char* const stack = reinterpret_cast<char*>(lds);           // [stack_rows][blockDim] x 16 B
    const uint32_t step = blockDim.x * (uint32_t)sizeof(uint4);  // next slot of the same lane
    const uint32_t base = tid * (uint32_t)sizeof(uint4);
    uint32_t* const wscratch = reinterpret_cast<uint32_t*>(lds + (size_t)a.stack_rows * blockDim.x) +
                               wave * 64u;                      // 64 u32 per wave
    const uint32_t count = *a.item_count;

    // lane state: st = {f, avail, ld, rd} of the current row; the lane's open
    // rows are the LDS slots [bot, top)
    uint4 st = make_uint4(0, 0, 0, 0);
    uint32_t mask = 0, solve = 0, ws = 0;
    uint32_t bot = base, top = base;
    bool live = false;
    uint64_t pnodes = 0;  // placements of the current piece
    sol_t psols = 0;      // solutions of the current piece
    uint32_t acc_solve = 0;
    uint64_t acc_sols = 0, acc_nodes = 0, exec_nodes = 0;

    bool queue_open = true;
    uint64_t pass = 0;  // wave-uniform: passes run so far

    for (;;) {
        // ---- pass budget: spill every open row as an item, then leave
        if (a.budget && pass >= a.budget) {
            if (live) {
                // open rows: the current one if it has candidates, then the stack
                const uint32_t cur = st.x != 0 ? 1u : 0u;
                const uint32_t nitems = cur + (top - bot) / step;
                const uint32_t o = nitems ? atomicAdd(a.spill_count, nitems) : 0u;
                if (nitems && (uint64_t)o + nitems > a.spill_cap) {
                    atomicOr(a.err, 4u);
                } else if (nitems) {
                    const uint32_t n = 31u - __clz(mask + 1u);
                    Item it;
                    it.solve = solve;
                    it.pad0 = it.pad1 = 0;
                    uint4 e = st;
                    uint32_t sa = bot;
                    for (uint32_t j = 0; j < nitems; ++j) {
                        if (j || !cur) {
                            e = *reinterpret_cast<const uint4*>(stack + sa);
                            sa += step;
                        }
                        it.f = e.x;
                        it.cols = ~e.y & mask;
                        it.ld = e.z & mask;
                        it.rd = e.w;
                        it.meta = item_meta(n, (uint32_t)__popc(it.cols), ws);
                        a.spill[o + j] = it;
                    }
                }
                exec_nodes += pnodes;
                acc_nodes += pnodes << ws;
                acc_sols += (uint64_t)psols << ws;
                live = false;
            }
            break;
        }

        // ---- refill idle lanes: the global queue first, then in-wave donation
        uint64_t idle = __ballot(!live);
        if (idle) {
            const uint32_t nidle = __popcll(idle);
            if (queue_open) {
                uint32_t qbase = 0;
                if (lane == 0) qbase = atomicAdd(a.head, nidle);
                qbase = __shfl(qbase, 0, 64);
                const uint32_t got = qbase >= count ? 0u : min(nidle, count - qbase);
                if (got < nidle) queue_open = false;
                const uint32_t r = lane_rank(idle);
                if (!live && r < got) {
                    const Item it = a.items[qbase + r];
                    if (it.solve != acc_solve) {
                        flush_acc(a, acc_solve, acc_sols, acc_nodes);
                        acc_solve = it.solve;
                    }
                    mask = (1u << meta_n(it.meta)) - 1u;
                    const uint32_t avail = ~it.cols & mask;
                    st = make_uint4(it.f, avail, it.ld, it.rd);
                    solve = it.solve;
                    ws = meta_wshift(it.meta);
                    bot = top = base;
                    pnodes = 0;
                    psols = 0;
                    live = true;
                }
                idle = __ballot(!live);
            }
            if (idle && a.donate) {
                const bool is_donor = live && top != bot;
                const uint64_t donors = __ballot(is_donor);
                const uint32_t k = min((uint32_t)__popcll(idle), (uint32_t)__popcll(donors));
                if (k) {
                    const uint32_t dr = lane_rank(donors);
                    if (is_donor && dr < k) wscratch[dr] = lane;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    const uint32_t ir = lane_rank(idle);
                    const bool recv = !live && ir < k;
                    const uint32_t dl = recv ? wscratch[ir] : lane;
                    const uint32_t d_bot = __shfl(bot, dl, 64);
                    const uint32_t d_mask = __shfl(mask, dl, 64);
                    const uint32_t d_solve = __shfl(solve, dl, 64);
                    const uint32_t d_ws = __shfl(ws, dl, 64);
                    if (recv) {
                        // the donor's shallowest open row, straight from its LDS slot
                        st = *reinterpret_cast<const uint4*>(stack + d_bot);
                        if (d_solve != acc_solve) {
                            flush_acc(a, acc_solve, acc_sols, acc_nodes);
                            acc_solve = d_solve;
                        }
                        mask = d_mask;
                        solve = d_solve;
                        ws = d_ws;
                        bot = top = base;
                        pnodes = 0;
                        psols = 0;
                        live = true;
                    }
                    if (is_donor && dr < k) bot += step;
                    // keep the scratch stable until every receiver has read it
                    __builtin_amdgcn_wave_barrier();
                }
            }
        }
        const uint64_t live_mask = __ballot(live);
        if (live_mask == 0 && !queue_open) break;

        // ---- search: live lanes place one queen per pass (the K1 step) until
        // their piece ends, or until enough lanes idle to make a refill worth
        // it (only while a refill can find work), or the pass budget is hit
        const uint32_t nlive = __popcll(live_mask);
        const bool can_refill = queue_open || (a.donate && __ballot(live && top != bot) != 0);
        uint32_t floor_live = 0;
        uint64_t limit = a.budget ? a.budget - pass : 0xffffffffffffull;
        if (can_refill) {
            if (nlive > 64u - a.refill_idle)
                floor_live = 64u - a.refill_idle;  // leave once refill_idle lanes idle
            else
                limit = limit < 256u ? limit : 256u;  // idle lanes already: retry soon
        }
        // thresholds are wave-uniform: keep them in scalar registers
        floor_live = __builtin_amdgcn_readfirstlane(floor_live);
        const uint32_t lim32 = __builtin_amdgcn_readfirstlane(
            (uint32_t)(limit < 0x7fffffffull ? limit : 0x7fffffffull));
        const uint64_t pn0 = pnodes;
        if (live) {
            uint32_t ran = 0;  // passes of this phase (uniform among running lanes)
            for (;;) {
                bool fin = false;
                if (st.x == 0) {  // row exhausted: resume the nearest open row
                    if (top == bot) {
                        fin = true;  // the piece is done
                    } else {
                        top -= step;
                        st = *reinterpret_cast<const uint4*>(stack + top);
                    }
                }
                if (!fin) {
                    const uint32_t b = st.x & (0u - st.x);
                    st.x ^= b;
                    const uint32_t a2 = st.y ^ b;
                    const uint32_t l2 = (st.z + b) << 1;  // b is not in ld: + is |
                    const uint32_t r2 = (st.w | b) >> 1;
                    const uint32_t f2 = a2 & ~l2 & ~r2;
                    psols += (a2 == 0);
                    const bool go = f2 != 0;
                    if (go && st.x != 0) {
                        *reinterpret_cast<uint4*>(stack + top) = st;
                        top += step;
                    }
                    st.x = go ? f2 : st.x;
                    st.y = go ? a2 : st.y;
                    st.z = go ? l2 : st.z;
                    st.w = go ? r2 : st.w;
                    ++pnodes;
                }
                ++ran;
                // one exit: finished lanes, or everyone when enough lanes are
                // done or the phase limit is reached
                const uint32_t running = __popcll(__ballot(!fin));
                if (fin || running <= floor_live || ran >= lim32) {
                    live = !fin;
                    break;
                }
            }
            if (!live) {  // account the finished piece
                exec_nodes += pnodes;
                acc_nodes += pnodes << ws;
                acc_sols += (uint64_t)psols << ws;
            }
        }
        // the phase length: the most placements any lane made in it
        uint64_t rmax = pnodes - pn0;
        for (int off = 32; off > 0; off >>= 1) {
            const uint64_t o = __shfl_xor(rmax, off, 64);
            rmax = o > rmax ? o : rmax;
        }
        rmax = rmax ? rmax : 1;  // a phase that only closed pieces still counts
        pass += rmax;
    }

    add_per_solve(a.solve_sols, acc_solve, acc_sols, lane);
    add_per_solve(a.solve_nodes, acc_solve, acc_nodes, lane);
    exec_nodes = wave_sum64(exec_nodes);
    if (lane == 0 && exec_nodes)
        atomicAdd((unsigned long long*)a.exec_nodes, (unsigned long long)exec_nodes);
}

// ---------------------------------------------------------------------------
// small helpers (graph-safe clears instead of memset nodes)
// ---------------------------------------------------------------------------
__global__ void k2_fill_u32_kernel(uint32_t* p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k2_fill_u64_kernel(uint64_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

hipError_t k2_fill_u32(hipStream_t s, uint32_t* p, uint32_t n, uint32_t v) {
    const uint32_t blocks = n ? min(1024u, (n + 255u) / 256u) : 1u;
    hipLaunchKernelGGL(k2_fill_u32_kernel, dim3(blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}
hipError_t k2_fill_u64(hipStream_t s, uint64_t* p, uint32_t n) {
    const uint32_t blocks = n ? min(1024u, (n + 255u) / 256u) : 1u;
    hipLaunchKernelGGL(k2_fill_u64_kernel, dim3(blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t k2_expand(hipStream_t s, uint32_t grid, const Item* in, const uint32_t* in_count,
                     Item* out, uint32_t* out_count, uint32_t out_cap, const uint8_t* target_row,
                     uint64_t* solve_nodes, uint64_t* exec_nodes, uint32_t* err, const float* pred,
                     float thresh) {
    hipLaunchKernelGGL(k2_expand_kernel, dim3(grid), dim3(256), 0, s, in, in_count, out, out_count,
                       out_cap, target_row, solve_nodes, exec_nodes, err, pred, thresh);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// largest-predicted-first order: counting sort on 64 buckets of half a log2
// ---------------------------------------------------------------------------
constexpr uint32_t kBuckets = 64;

__device__ __forceinline__ uint32_t pred_bucket(float p) {
    // bucket 0 holds the largest predictions (log2 sizes 31.5 and up)
    const float b = (32.0f - p) * 2.0f;
    return b <= 0.f ? 0u : (b >= (float)(kBuckets - 1) ? kBuckets - 1 : (uint32_t)b);
}

__global__ void k2_bucket_hist_kernel(const float* __restrict__ pred, const uint32_t* __restrict__ count,
                                      uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kBuckets];
    for (uint32_t i = threadIdx.x; i < kBuckets; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const uint32_t c = *count;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < c; i += gridDim.x * blockDim.x)
        atomicAdd(&h[pred_bucket(pred[i])], 1u);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kBuckets; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

__global__ void k2_bucket_scan_kernel(uint32_t* __restrict__ hist) {
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (uint32_t i = 0; i < kBuckets; ++i) {
            const uint32_t v = hist[i];
            hist[i] = acc;
            acc += v;
        }
    }
}

__global__ void k2_bucket_scatter_kernel(const Item* __restrict__ in, const float* __restrict__ pred,
                                         const uint32_t* __restrict__ count, uint32_t* __restrict__ cursor,
                                         Item* __restrict__ out) {
    const uint32_t c = *count;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < c; i += gridDim.x * blockDim.x) {
        const uint32_t o = atomicAdd(&cursor[pred_bucket(pred[i])], 1u);
        out[o] = in[i];
    }
}

hipError_t k2_order_by_prediction(hipStream_t s, const Item* in, const float* pred, const uint32_t* count,
                                  uint32_t max_count, uint32_t* hist, Item* out) {
    const uint32_t blocks = std::max(1u, std::min(2048u, (max_count + 255u) / 256u));
    hipLaunchKernelGGL(k2_fill_u32_kernel, dim3(1), dim3(64), 0, s, hist, kBuckets, 0u);
    hipLaunchKernelGGL(k2_bucket_hist_kernel, dim3(blocks), dim3(256), 0, s, pred, count, hist);
    hipLaunchKernelGGL(k2_bucket_scan_kernel, dim3(1), dim3(64), 0, s, hist);
    hipLaunchKernelGGL(k2_bucket_scatter_kernel, dim3(blocks), dim3(256), 0, s, in, pred, count, hist, out);
    return hipGetLastError();
}

// Items a budget-ended round never handed out: items[head .. count) move to
// the end of the spill queue so the next round sees them.
__global__ void k2_append_leftover_kernel(const Item* __restrict__ items,
                                          const uint32_t* __restrict__ count,
                                          const uint32_t* __restrict__ head, Item* __restrict__ out,
                                          uint32_t* __restrict__ out_count, uint32_t out_cap,
                                          uint32_t* __restrict__ err) {
    const uint32_t c = *count;
    const uint32_t h = min(*head, c);
    const uint32_t left = c - h;
    const uint32_t o = *out_count;  // stable: the drain round has ended
    if (left == 0) return;
    if ((uint64_t)o + left > out_cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(err, 8u);
        return;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < left; i += gridDim.x * blockDim.x)
        out[o + i] = items[h + i];
}

__global__ void k2_bump_count_kernel(const uint32_t* __restrict__ count,
                                     const uint32_t* __restrict__ head, uint32_t* __restrict__ out_count,
                                     uint32_t out_cap) {
    const uint32_t c = *count;
    const uint32_t h = min(*head, c);
    if ((uint64_t)*out_count + (c - h) <= out_cap) *out_count += c - h;
}

hipError_t k2_append_leftover(hipStream_t s, const Item* items, const uint32_t* count,
                              const uint32_t* head, Item* out, uint32_t* out_count, uint32_t out_cap,
                              uint32_t* err) {
    hipLaunchKernelGGL(k2_append_leftover_kernel, dim3(1024), dim3(256), 0, s, items, count, head, out,
                       out_count, out_cap, err);
    hipLaunchKernelGGL(k2_bump_count_kernel, dim3(1), dim3(1), 0, s, count, head, out_count, out_cap);
    return hipGetLastError();
}

size_t k2_lds_bytes(uint32_t stack_rows, uint32_t threads) {
    return (size_t)stack_rows * threads * sizeof(uint4) + (size_t)threads * sizeof(uint32_t);
}

hipError_t k2_drain(hipStream_t s, uint32_t blocks, uint32_t threads, const K2DrainParams& p) {
    DrainArgs a;
    a.items = p.items;
    a.item_count = p.item_count;
    a.head = p.head;
    a.spill = p.spill;
    a.spill_count = p.spill_count;
    a.spill_cap = p.spill_cap;
    a.solve_sols = p.solve_sols;
    a.solve_nodes = p.solve_nodes;
    a.exec_nodes = p.exec_nodes;
    a.err = p.err;
    a.stack_rows = p.stack_rows;
    a.refill_idle = p.refill_idle;
    a.donate = p.donate;
    a.pad = 0;
    a.budget = p.budget;
    if (p.wide)
        hipLaunchKernelGGL(k2_drain_kernel<true>, dim3(blocks), dim3(threads),
                           k2_lds_bytes(p.stack_rows, threads), s, a);
    else
        hipLaunchKernelGGL(k2_drain_kernel<false>, dim3(blocks), dim3(threads),
                           k2_lds_bytes(p.stack_rows, threads), s, a);
    return hipGetLastError();
}

int k2_max_blocks_per_cu(uint32_t threads, size_t lds_bytes, int* out) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k2_drain_kernel<true>, (int)threads, lds_bytes) !=
        hipSuccess) {
        (void)hipGetLastError();
        return NQM_EDEVICE;
    }
    *out = nb;
    return NQM_OK;
}

}  // namespace nqm
