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
//    shallowest open row, i.e. the largest untried subtree), which the
//    receiver reads straight from the donor's slot, and moves its other
//    open rows down one slot. Solves of different n mix freely: each lane
//    holds its own column mask. Counts are folded per lane and flushed to
//    per-solve u64 counters when the lane switches solve or exits.
//
//  * Optional pass budget: when a wave has run `budget` passes, each lane
//    with an unfinished piece parks it (current row + open rows) as one
//    continuation record, hands back the queue items its wave holds, and the
//    wave exits. The host relaunches (restart): a later round takes queue
//    items first (so new work joins between rounds), then resumes the parked
//    pieces where they stopped.

#include <algorithm>
#include <type_traits>

#include "nqm_internal.h"
#include "nqm_items.h"

namespace nqm {


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

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor(v, off, 64));
    return __builtin_amdgcn_readfirstlane(v);
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
// Block b owns the contiguous chunk [b*c, (b+1)*c) of the input, c = ceil(cnt
// / gridDim). The count pass stores each block's number of children and its
// placements (a per-block record, no atomics); one block sums the level
// size. The write pass places each block's children after those of the
// blocks ranked before it, and one block then adds the placements to the
// per-solve counters. Same-address atomics, which serialize across the
// chip, are left to blocks that hold items of several solves.
//
// Rank = block index keeps input order (the learned stage's largest-first
// order carries over to the children). Rank = the bit-reversed block index
// interleaves the input's regions: a drain then works on many solves and on
// big and small subtrees at once, instead of sweeping one solve's items at a
// time (all lanes updating that solve's counters together).

// Which input items expand, and into how many live children.
struct ExpandRule {
    const uint8_t* target_row;
    const float* pred;
    float thresh;
    __device__ __forceinline__ bool expands(const Item& it, uint32_t i) const {
        const uint32_t n = meta_n(it.meta), row = meta_row(it.meta);
        // by prediction (split the large subtrees, never the last rows) or
        // up to the solve's target row
        return pred ? (pred[i] > thresh && row + 2 < n) : (row < target_row[it.solve]);
    }
};

__device__ __forceinline__ uint32_t live_children(const Item& it) {
    const uint32_t mask = (1u << meta_n(it.meta)) - 1u;
    uint32_t f = it.f, nch = 0;
    while (f) {
        const uint32_t b = f & (0u - f);
        f ^= b;
        nch += (~((it.cols | b) | ((it.ld | b) << 1) | ((it.rd | b) >> 1)) & mask) != 0u;
    }
    return nch;
}

// The candidates of an item whose child has a free cell.
__device__ __forceinline__ uint32_t live_mask(const Item& it) {
    const uint32_t mask = (1u << meta_n(it.meta)) - 1u;
    uint32_t f = it.f, live = 0;
    while (f) {
        const uint32_t b = f & (0u - f);
        f ^= b;
        live |= (~((it.cols | b) | ((it.ld | b) << 1) | ((it.rd | b) >> 1)) & mask) != 0u ? b : 0u;
    }
    return live;
}

// The child of an item for candidate b.
__device__ __forceinline__ Item child_item(const Item& it, uint32_t b) {
    const uint32_t n = meta_n(it.meta);
    const uint32_t mask = (1u << n) - 1u;
    Item c;
    c.cols = it.cols | b;
    c.ld = ((it.ld | b) << 1) & mask;
    c.rd = (it.rd | b) >> 1;
    c.f = ~(c.cols | c.ld | c.rd) & mask;
    c.solve = it.solve;
    c.meta = item_meta(n, meta_row(it.meta) + 1u, meta_wshift(it.meta));
    c.pad0 = c.pad1 = 0;
    return c;
}

// Block-wide exclusive scan of one u32 per thread (blockDim = 256); *total
// receives the block sum. Ends with the scratch free for reuse.
__device__ __forceinline__ uint32_t block_scan256(uint32_t x, uint32_t* scratch, uint32_t* total) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t incl = wave_incl_scan(x, lane);
    if (lane == 63u) scratch[wave] = incl;
    __syncthreads();
    uint32_t before = 0, sum = 0;
#pragma unroll
    for (uint32_t w = 0; w < 4u; ++w) {
        const uint32_t v = scratch[w];
        before += w < wave ? v : 0u;
        sum += v;
    }
    __syncthreads();
    *total = sum;
    return before + incl - x;
}

// Block-wide sum of one u64 per thread (blockDim = 256), returned to every
// thread. Ends with the scratch free for reuse.
__device__ __forceinline__ uint64_t block_sum256_u64(uint64_t x, uint64_t* scratch) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    x = wave_sum64(x);
    if (lane == 0u) scratch[wave] = x;
    __syncthreads();
    const uint64_t sum = scratch[0] + scratch[1] + scratch[2] + scratch[3];
    __syncthreads();
    return sum;
}

// Per-block record of one level: children, and the block's placements when
// all its expanded items belong to one solve (solve == kMixed: the write
// pass adds them itself, per wave).
struct ExpandPart {
    uint32_t children, solve;
    uint64_t nodes, exec;  // weighted and executed placements
    uint64_t pad;
};
static_assert(sizeof(ExpandPart) == 32, "kExpandPartBytes assumes 32-byte records");
constexpr uint32_t kMixed = 0xffffffffu;

__device__ __forceinline__ void expand_chunk(uint32_t cnt, uint32_t* c0, uint32_t* c1) {
    const uint32_t chunk = (cnt + gridDim.x - 1u) / gridDim.x;
    *c0 = min(cnt, blockIdx.x * chunk);
    *c1 = min(cnt, *c0 + chunk);
}

__global__ void __launch_bounds__(256)
k2_expand_count_kernel(const Item* __restrict__ in, const uint32_t* __restrict__ in_count,
                       ExpandPart* __restrict__ part, const uint8_t* __restrict__ target_row,
                       const float* __restrict__ pred, const float* __restrict__ thresh_p) {
    __shared__ uint32_t scratch[4];
    __shared__ uint64_t scratch64[4];
    __shared__ uint32_t s_mixed;
    const ExpandRule rule{target_row, pred, pred ? *thresh_p : 0.f};
    uint32_t c0, c1;
    expand_chunk(*in_count, &c0, &c1);
    if (threadIdx.x == 0) s_mixed = 0;
    __syncthreads();
    const uint32_t solve0 = c0 < c1 ? in[c0].solve : 0u;
    uint32_t mine = 0;
    uint64_t nodes = 0, exec = 0;
    bool mixed = false;
    for (uint32_t i = c0 + threadIdx.x; i < c1; i += 256u) {
        const Item it = in[i];
        if (rule.expands(it, i)) {
            mine += live_children(it);
            const uint32_t placed = (uint32_t)__popc(it.f);
            nodes += (uint64_t)placed << meta_wshift(it.meta);
            exec += placed;
            mixed |= it.solve != solve0;
        } else {
            mine += 1u;
        }
    }
    if (mixed) s_mixed = 1u;  // benign race: every writer stores 1
    uint32_t total = 0;
    (void)block_scan256(mine, scratch, &total);
    nodes = block_sum256_u64(nodes, scratch64);
    exec = block_sum256_u64(exec, scratch64);
    if (threadIdx.x == 0) {
        ExpandPart r;
        r.children = total;
        r.solve = s_mixed ? kMixed : solve0;
        r.nodes = nodes;
        r.exec = exec;
        r.pad = 0;
        part[blockIdx.x] = r;
    }
}

// One block: *out_count = the level size (saturated at 2^32 - 1, which the
// host rejects).
__global__ void __launch_bounds__(256)
k2_expand_total_kernel(const ExpandPart* __restrict__ part, uint32_t nblocks, uint32_t* __restrict__ out_count) {
    __shared__ uint64_t scratch64[4];
    uint64_t mine = 0;
    for (uint32_t j = threadIdx.x; j < nblocks; j += 256u) mine += part[j].children;
    const uint64_t total = block_sum256_u64(mine, scratch64);
    if (threadIdx.x == 0) *out_count = (uint32_t)min(total, (uint64_t)0xffffffffu);
}

__global__ void __launch_bounds__(256)
k2_expand_write_kernel(const Item* __restrict__ in, const uint32_t* __restrict__ in_count,
                       const ExpandPart* __restrict__ part, uint32_t interleave, Item* __restrict__ out,
                       uint32_t out_cap, const uint8_t* __restrict__ target_row, uint64_t* __restrict__ solve_nodes,
                       uint64_t* __restrict__ exec_nodes, uint32_t* __restrict__ err,
                       const float* __restrict__ pred, const float* __restrict__ thresh_p) {
    __shared__ uint32_t scratch[4];
    __shared__ Item s_item[256];
    __shared__ uint32_t s_live[256];  // live-child mask, 0 for an item copied as is
    __shared__ uint32_t s_off[256];   // first output of the item within the round
    __shared__ uint8_t s_map[256u * 31u];  // round output -> item (at most 31 children each)
    const ExpandRule rule{target_row, pred, pred ? *thresh_p : 0.f};
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t c0, c1;
    expand_chunk(*in_count, &c0, &c1);
    // this block's first output: the children of the blocks ranked before it
    uint32_t before = 0;
    if (interleave) {
        const uint32_t key = __brev(blockIdx.x);
        for (uint32_t j = threadIdx.x; j < gridDim.x; j += 256u)
            before += __brev(j) < key ? part[j].children : 0u;
    } else {
        for (uint32_t j = threadIdx.x; j < blockIdx.x; j += 256u) before += part[j].children;
    }
    uint32_t obase = 0;
    (void)block_scan256(before, scratch, &obase);
    const ExpandPart mine_part = part[blockIdx.x];
    const bool fits = (uint64_t)obase + mine_part.children <= out_cap;
    if (!fits && threadIdx.x == 0) atomicOr(err, 2u);
    const bool own_nodes = mine_part.solve == kMixed;  // else added by the commit kernel

    uint32_t acc_solve = 0;
    uint64_t acc_nodes = 0, acc_exec = 0;
    uint32_t o = obase;
    for (uint32_t b0 = c0; b0 < c1; b0 += 256u) {  // block-uniform trip count
        const uint32_t i = b0 + threadIdx.x;
        Item it{};
        bool expand = false;
        uint32_t live = 0, nch = 0;
        if (i < c1) {
            it = in[i];
            expand = rule.expands(it, i);
            live = expand ? live_mask(it) : 0u;
            nch = expand ? (uint32_t)__popc(live) : 1u;
        }
        uint32_t step_total = 0;
        const uint32_t wl = block_scan256(nch, scratch, &step_total);
        if (expand && own_nodes) {
            const uint32_t placed = (uint32_t)__popc(it.f);
            if (it.solve != acc_solve && acc_nodes) {
                atomicAdd((unsigned long long*)&solve_nodes[acc_solve], (unsigned long long)acc_nodes);
                acc_nodes = 0;
            }
            acc_solve = it.solve;
            acc_nodes += (uint64_t)placed << meta_wshift(it.meta);
            acc_exec += placed;
        }
        if (fits) {  // block-uniform: stage the round in LDS, then thread k writes outputs k, k+256, ...
            s_item[threadIdx.x] = it;
            s_live[threadIdx.x] = live;
            s_off[threadIdx.x] = wl;
            for (uint32_t c = 0; c < nch; ++c) s_map[wl + c] = (uint8_t)threadIdx.x;
            __syncthreads();
            for (uint32_t k = threadIdx.x; k < step_total; k += 256u) {
                const uint32_t p = s_map[k];
                uint32_t lm = s_live[p];
                Item c = s_item[p];
                if (lm) {  // the (k - off)-th live candidate
                    for (uint32_t q = k - s_off[p]; q; --q) lm &= lm - 1u;
                    c = child_item(c, lm & (0u - lm));
                }
                out[o + k] = c;
            }
            __syncthreads();
        }
        o += step_total;
    }
    if (own_nodes) {  // block-uniform
        add_per_solve(solve_nodes, acc_solve, acc_nodes, lane);
        const uint64_t wexec = wave_sum64(acc_exec);
        if (lane == 0 && wexec) atomicAdd((unsigned long long*)exec_nodes, (unsigned long long)wexec);
    }
}

// One block: the placements of the single-solve blocks of a written level.
// Consecutive blocks mostly share a solve, so each thread sums a run of
// blocks per solve and the waves add them up (a few atomics per level).
__global__ void __launch_bounds__(256)
k2_expand_commit_kernel(const ExpandPart* __restrict__ part, uint32_t nblocks,
                        uint64_t* __restrict__ solve_nodes, uint64_t* __restrict__ exec_nodes) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t per = (nblocks + 255u) / 256u;
    const uint32_t j0 = min(nblocks, threadIdx.x * per), j1 = min(nblocks, j0 + per);
    uint32_t acc_solve = 0;
    uint64_t acc_nodes = 0, acc_exec = 0;
    for (uint32_t j = j0; j < j1; ++j) {
        const ExpandPart r = part[j];
        if (r.solve == kMixed || r.exec == 0) continue;
        if (r.solve != acc_solve && acc_nodes) {
            atomicAdd((unsigned long long*)&solve_nodes[acc_solve], (unsigned long long)acc_nodes);
            acc_nodes = 0;
        }
        acc_solve = r.solve;
        acc_nodes += r.nodes;
        acc_exec += r.exec;
    }
    add_per_solve(solve_nodes, acc_solve, acc_nodes, lane);
    const uint64_t wexec = wave_sum64(acc_exec);
    if (lane == 0 && wexec) atomicAdd((unsigned long long*)exec_nodes, (unsigned long long)wexec);
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
    uint32_t next_refill;  // a phase ends once this many lanes started their next item
    uint32_t donate;
    uint32_t cont_cap;
    uint64_t budget;  // 0: drain everything
    uint32_t items_first;
    uint32_t nseg;  // queue segments, each with its own head (head[s * kHeadStride])
    const Cont* cont_in;
    const uint32_t* cont_in_count;
    uint32_t* cont_head;
    Cont* cont_out;
    uint32_t* cont_out_count;
    uint64_t* wave_part;  // nullable: per-wave totals {solve, sols, nodes, exec}
};

// Lanes that take a piece of another solve (sw) hand over what they counted
// for the old one. Lanes handing over the same solve are summed first: one
// atomic per solve and wave, not per lane (lanes of a wave usually change
// solve together). All lanes must call.
__device__ __forceinline__ void switch_solve(const DrainArgs& a, bool sw, uint32_t new_solve, uint32_t& acc_solve,
                                             uint64_t& sols, uint64_t& nodes, uint32_t lane) {
    bool pend = sw && (sols | nodes) != 0;
    for (uint64_t m = __ballot(pend); m; m = __ballot(pend)) {
        const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1u;
        const uint32_t s = __shfl(acc_solve, leader, 64);
        const bool mine = pend && acc_solve == s;
        const uint64_t ts = wave_sum64(mine ? sols : 0u), tn = wave_sum64(mine ? nodes : 0u);
        if (lane == leader) {
            if (ts) atomicAdd((unsigned long long*)&a.solve_sols[s], (unsigned long long)ts);
            if (tn) atomicAdd((unsigned long long*)&a.solve_nodes[s], (unsigned long long)tn);
        }
        if (mine) {
            sols = 0;
            nodes = 0;
            pend = false;
        }
    }
    if (sw) acc_solve = new_solve;
}

// LDS of one drain workgroup of T lanes (T = 64, 128 or 256), rows of 16 T
// bytes:
//   row 0                  the donation scratch (64 u32 per wave); a lane
//                          whose stack is empty reads its 16 bytes at 16 t
//                          here when it pops, and ignores them;
//   row j + 1              slot j of every lane: 16 bytes {f, avail, ld, rd}
//                          at 16 t + 16 T (j + 1). A row is 16 T consecutive
//                          bytes, so the ds_read_b128 / ds_write_b128 of a
//                          wave reads or writes each 4-byte bank once per 16
//                          lanes whatever the stack depth of each lane (the
//                          row stride is a multiple of the 256 bytes the banks
//                          span). One b128 read of a wave costs 4 LDS cycles
//                          on gfx950 against 7.3 for the two b64 halves of
//                          ds_read2st64_b64 (measured).
// A lane's stack register is q = 16 t + 16 T k with k open rows. A push
// stores at q + 16 T (the offset field of the instruction) and adds 16 T; a
// pop subtracts 16 T and loads at q + 16 T. The subtraction borrows exactly
// when the stack is empty (16 t < 16 T), so one v_sub_co_u32 both moves the
// pointer and says that the piece is done; the load then reads 16 t (the
// address sum wraps modulo 2^32), the lane's own slot of row 0.
template <uint32_t T>
struct Stack {
    static constexpr uint32_t kRow = 16u * T;  // bytes from one slot of a lane to its next
    static __device__ __forceinline__ uint32_t base(uint32_t t) { return 16u * t; }
    static size_t bytes(uint32_t rows) { return (size_t)(rows + 1u) * kRow; }
};
static_assert(64u * sizeof(uint32_t) <= 16u * 64u, "a wave's donation scratch fits its lanes' row-0 slots");

// Stack slots are addressed by their LDS byte offset; the dynamic LDS of the
// drain is the only LDS of the kernel, so offset == address.
typedef uint32_t lds_v4u_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) lds_v4u_t lds_v4u;
typedef __attribute__((address_space(3))) uint32_t lds_u32;
__device__ __forceinline__ uint4 slot_load(uint32_t addr) {
    const lds_v4u_t v = *reinterpret_cast<const lds_v4u*>((size_t)addr);
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void slot_store(uint32_t addr, uint4 s) {
    *reinterpret_cast<lds_v4u*>((size_t)addr) = lds_v4u_t{s.x, s.y, s.z, s.w};
}

#ifndef NQM_PASS_UNROLL
#define NQM_PASS_UNROLL 8
#endif
constexpr uint32_t kPassUnroll = NQM_PASS_UNROLL;  // passes between two votes
constexpr uint32_t kPhaseCap = 1u << 26;     // passes per phase: u32 counters cannot wrap
constexpr uint32_t kGrabMax = kGrabItems;    // queue items one wave takes per atomic

// One search phase, entered by the lanes that hold a piece, written as one
// asm block because it manages the execution mask itself: a lane whose piece
// is done leaves the mask at once (it issues nothing more), and the phase
// ends at a wave-uniform check every 8 passes, once at most floor_live
// lanes are left, at most nfloor lanes still hold an unstarted next item
// (signed: -1 never ends it) or lim32 passes ran. In C++ the compiler's lowering of this
// control flow spent 4-7 extra instructions per pass on copies and mask
// merges. Here a pass is 13 VALU (2 for the pop, 11 for the step), 5 SALU
// (with a wave-counted solution sum), one branch, one s_waitcnt and at most
// one ds_read_b128 and one ds_write_b128 (measured at N=18: drain 15.4 ms
// against 17.4 ms for the compiled loop with two b64 halves per slot).
//
// The drain is bound by the latency of this chain, not by LDS or VALU
// throughput: a pop's b128 read returns after about 60 cycles unloaded and
// 130 at 16 waves per CU, and the pass cannot place before it returns. So
// the block keeps the chain short: the pop's mask comes from v_cmpx (no
// SALU wait on a VALU result), the finished-piece bookkeeping sits in a
// branch a wave rarely takes, and two register sets hold the current row
// alternately (a pass reads its row from one and writes the child into the
// other), so the push still has the row intact and the child's ops need not
// wait for it.
//
// Registers: v[88:91] and v[92:95] = {f, avail, ld, rd} of the current row
// (the b128 tuple of the pop and the push), alternating per pass; q = the
// lane's stack register (layout above); alive = the lanes still placing. A
// pass:
//   pop    lanes whose row has no candidate left: q -= 16 T and load the
//          slot at q + 16 T. The borrow marks a finished piece: a lane in
//          hasnext (it holds its next queue item, of its own solve and
//          weight, in v[96:99]) starts that item in this pass (q back to its
//          base, the bit leaves hasnext); any other leaves alive, and its
//          placements, the pass index, are added to nodes;
//   step   b = f & -f; a solution when avail == b (every column taken);
//          lanes with candidates left push {f ^ b, avail, ld, rd} at
//          q + 16 T and move q up; the child into the other set:
//          avail ^ b, (ld + b) << 1 (b is not in ld, so + is |),
//          (rd | b) >> 1 and its free set avail & ~ld & ~rd in one v_bitop3
//          (truth table 0x10).
// Always entering the child keeps the step free of selects: a child
// without a free cell is left again by the next pass's pop, which also
// places the parent's next candidate, so every pass is one placement.
// Loads complete inside the block (s_waitcnt before the first use), and
// the block ends with the execution mask it started with. The carry-out of
// v_subrev_co_u32 is zero for lanes off the mask, so it is the borrow mask
// as it is.
//
// kWaveSols (every lane of one solve and weight): *psols and *pnodes are
// the wave's solutions and placements, summed from lane masks in scalar
// registers; else each lane's own. *ran = passes of the phase.
#define NQM_K2_POP(U, CX, CA, CL, CR, CT)                  \
    "v_cmpx_eq_u32 vcc, 0, " CX "\n"                       \
    "v_subrev_co_u32 %[q], vcc, %[row], %[q]\n"            \
    "ds_read_b128 " CT ", %[q] offset:%[row]\n"            \
    "s_cbranch_vccz .Lk2_nd" #U "_%=\n"                    \
    "s_and_b64 exec, vcc, %[hasnext]\n"                    \
    "s_cbranch_execz .Lk2_nn" #U "_%=\n"                   \
    "s_waitcnt lgkmcnt(0)\n"                               \
    "v_mov_b32 " CX ", v96\n"                              \
    "v_mov_b32 " CA ", v97\n"                              \
    "v_mov_b32 " CL ", v98\n"                              \
    "v_mov_b32 " CR ", v99\n"                              \
    "v_add_u32 %[q], %[row], %[q]\n"                       \
    "s_andn2_b64 %[hasnext], %[hasnext], exec\n"           \
    "s_andn2_b64 vcc, vcc, exec\n"                         \
    ".Lk2_nn" #U "_%=:\n"                                  \
    "s_andn2_b64 %[alive], %[alive], vcc\n"                \
    "s_bcnt1_i32_b64 %[t], vcc\n"                          \
    "s_add_u32 %[t2], %[ran], " #U "\n"                    \
    "s_mul_i32 %[t], %[t], %[t2]\n"                        \
    "s_add_u32 %[nodes], %[nodes], %[t]\n"                 \
    ".Lk2_nd" #U "_%=:\n"                                  \
    "s_mov_b64 exec, %[alive]\n"                           \
    "s_waitcnt lgkmcnt(0)\n"
#define NQM_K2_STEP(CX, CA, CL, CR, CT, NX, NA, NL, NR, SOLCOUNT) \
    "v_sub_u32 %[n], 0, " CX "\n"                          \
    "v_and_b32 %[b], " CX ", %[n]\n"                       \
    "v_xor_b32 " NA ", " CA ", %[b]\n"                     \
    "v_or_b32 " NR ", " CR ", %[b]\n"                      \
    "v_add_lshl_u32 " NL ", %[b], " CL ", 1\n"             \
    "v_cmp_eq_u32 %[sm], " CA ", %[b]\n"                   \
    "v_cmpx_ne_u32 vcc, " CX ", %[b]\n"                    \
    "v_xor_b32 " CX ", " CX ", %[b]\n"                     \
    "ds_write_b128 %[q], " CT " offset:%[row]\n"           \
    "v_add_u32 %[q], %[row], %[q]\n"                       \
    "s_mov_b64 exec, %[alive]\n"                           \
    "v_lshrrev_b32 " NR ", 1, " NR "\n"                    \
    SOLCOUNT                                               \
    "v_bitop3_b32 " NX ", " NA ", " NL ", " NR " bitop3:0x10\n"
#define NQM_K2_SOLS_WAVE                                   \
    "s_bcnt1_i32_b64 %[t], %[sm]\n"                        \
    "s_add_u32 %[sols], %[sols], %[t]\n"
#define NQM_K2_SOLS_LANE                                   \
    "v_addc_co_u32 %[sols], vcc, 0, %[sols], %[sm]\n"
#define NQM_K2_PASS_A(U, SC) NQM_K2_POP(U, "v88", "v89", "v90", "v91", "v[88:91]") NQM_K2_STEP("v88", "v89", "v90", "v91", "v[88:91]", "v92", "v93", "v94", "v95", SC)
#define NQM_K2_PASS_B(U, SC) NQM_K2_POP(U, "v92", "v93", "v94", "v95", "v[92:95]") NQM_K2_STEP("v92", "v93", "v94", "v95", "v[92:95]", "v88", "v89", "v90", "v91", SC)
#define NQM_K2_PASS_WAVE(U) NQM_K2_PASS_##U(NQM_K2_SOLS_WAVE)
#define NQM_K2_PASS_LANE(U) NQM_K2_PASS_##U(NQM_K2_SOLS_LANE) "v_add_u32 %[lnodes], 1, %[lnodes]\n"
#define NQM_K2_PASS_0(SC) NQM_K2_PASS_A(0, SC)
#define NQM_K2_PASS_1(SC) NQM_K2_PASS_B(1, SC)
#define NQM_K2_PASS_2(SC) NQM_K2_PASS_A(2, SC)
#define NQM_K2_PASS_3(SC) NQM_K2_PASS_B(3, SC)
#define NQM_K2_PASS_4(SC) NQM_K2_PASS_A(4, SC)
#define NQM_K2_PASS_5(SC) NQM_K2_PASS_B(5, SC)
#define NQM_K2_PASS_6(SC) NQM_K2_PASS_A(6, SC)
#define NQM_K2_PASS_7(SC) NQM_K2_PASS_B(7, SC)
#define NQM_K2_PASS8(P) P(0) P(1) P(2) P(3) P(4) P(5) P(6) P(7)
#define NQM_K2_PHASE(P, END)                               \
    "s_mov_b64 %[live], exec\n"                            \
    "s_mov_b64 %[alive], exec\n"                           \
    "s_mov_b32 %[ran], 0\n"                                \
    ".Lk2_pass_%=:\n" NQM_K2_PASS8(P)                      \
    "s_add_u32 %[ran], %[ran], 8\n"                        \
    "s_bcnt1_i32_b64 %[t], exec\n"                         \
    "s_cmp_le_u32 %[t], %[floor]\n"                        \
    "s_cbranch_scc1 .Lk2_end_%=\n"                         \
    "s_bcnt1_i32_b64 %[t], %[hasnext]\n"                   \
    "s_cmp_le_i32 %[t], %[nfloor]\n"                       \
    "s_cbranch_scc1 .Lk2_end_%=\n"                         \
    "s_cmp_lt_u32 %[ran], %[lim]\n"                        \
    "s_cbranch_scc1 .Lk2_pass_%=\n"                        \
    ".Lk2_end_%=:\n" END                                   \
    "s_mov_b64 exec, %[live]\n"                            \
    "s_waitcnt lgkmcnt(0)\n"
static_assert(NQM_PASS_UNROLL == 8, "the asm phase votes every 8 passes");

typedef uint32_t v4u_t __attribute__((ext_vector_type(4)));

template <uint32_t T, bool kWaveSols>
__device__ __forceinline__ void search_phase(uint4& st, uint32_t& q, const uint4& nx, uint64_t& hasnext,
                                             int32_t nfloor, uint32_t floor_live, uint32_t lim32,
                                             uint32_t& psols, uint32_t& pnodes, uint32_t& ran) {
    v4u_t s4 = {st.x, st.y, st.z, st.w};
    v4u_t s5;  // the other register set (scratch outside the block)
    const v4u_t n4 = {nx.x, nx.y, nx.z, nx.w};
    uint64_t live, alive, sm;
    uint32_t b, n, t, t2;
    uint32_t sols = 0, nodes = 0;
    if (kWaveSols) {
        asm volatile("s_mov_b32 %[sols], 0\n"
                     "s_mov_b32 %[nodes], 0\n" NQM_K2_PHASE(NQM_K2_PASS_WAVE,
                                                            "s_bcnt1_i32_b64 %[t], exec\n"
                                                            "s_mul_i32 %[t], %[t], %[ran]\n"
                                                            "s_add_u32 %[nodes], %[nodes], %[t]\n")
                     : "+{v[88:91]}"(s4), "={v[92:95]}"(s5), [q] "+v"(q), [b] "=&v"(b), [n] "=&v"(n),
                       [live] "=&s"(live), [alive] "=&s"(alive), [sm] "=&s"(sm), [t] "=&s"(t), [t2] "=&s"(t2),
                       [ran] "=&s"(ran), [sols] "=&s"(sols), [nodes] "=&s"(nodes), [hasnext] "+s"(hasnext)
                     : [row] "i"(Stack<T>::kRow), [floor] "s"(floor_live), [lim] "s"(lim32), [nfloor] "s"(nfloor),
                       "{v[96:99]}"(n4)
                     : "vcc", "scc", "memory");
        psols = sols;
        pnodes = nodes;
    } else {
        // the finished-piece sum in nodes is unused here: lnodes counts
        // each lane's own placements
        uint32_t lsols = psols, lnodes = pnodes;
        asm volatile(NQM_K2_PHASE(NQM_K2_PASS_LANE, "")
                     : "+{v[88:91]}"(s4), "={v[92:95]}"(s5), [q] "+v"(q), [b] "=&v"(b), [n] "=&v"(n),
                       [live] "=&s"(live), [alive] "=&s"(alive), [sm] "=&s"(sm), [t] "=&s"(t), [t2] "=&s"(t2),
                       [ran] "=&s"(ran), [sols] "+v"(lsols), [nodes] "=&s"(nodes), [lnodes] "+v"(lnodes),
                       [hasnext] "+s"(hasnext)
                     : [row] "i"(Stack<T>::kRow), [floor] "s"(floor_live), [lim] "s"(lim32), [nfloor] "s"(nfloor),
                       "{v[96:99]}"(n4)
                     : "vcc", "scc", "memory");
        psols = lsols;
        pnodes = lnodes;
    }
    (void)s5;
    st = make_uint4(s4.x, s4.y, s4.z, s4.w);
}

template <uint32_t T>
__global__ void __launch_bounds__(T) k2_drain_kernel(DrainArgs a) {
    extern __shared__ uint4 lds[];
    constexpr uint32_t row = Stack<T>::kRow;  // next slot of the same lane
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    const uint32_t base = Stack<T>::base(tid);
    uint32_t* const wscratch = reinterpret_cast<uint32_t*>(lds) + wave * 64u;  // in row 0
    // the stack offsets below are LDS addresses: the dynamic LDS must start at 0
    if ((uint32_t)(size_t)(lds_u32*)lds != 0u) {
        if (tid == 0) atomicOr(a.err, 16u);
        return;
    }
    const uint32_t count = *a.item_count;
    const uint32_t waves = gridDim.x * (T / 64u);

    // lane state: st = {f, avail, ld, rd} of the current row; the open rows
    // are slots 0..k-1, k = (q - base) / row (slot j at base + row (j + 1)).
    // The lane holds a piece while st.x != 0 or q != base.
    uint4 st = make_uint4(0u, 0u, 0u, 0u);
    uint32_t mask = 0, solve = 0, ws = 0;
    uint32_t q = base;
    uint32_t acc_solve = 0;
    uint64_t acc_sols = 0, acc_nodes = 0, exec_nodes = 0;
    auto holds = [&]() { return st.x != 0u || q != base; };

    // wave-uniform queue state: items [lq, lq_end) are taken from the queue
    // and not yet started; rem is what the current segment held after the
    // last grab. The queue is cut into nseg segments with a head each; a
    // wave starts on the segment of its workgroup (workgroups go round-robin
    // to the XCDs) and moves on when it finds one used up, so the head
    // atomics of the chip spread over nseg addresses.
    const uint32_t nseg = a.nseg;
    uint32_t seg = blockIdx.x % nseg, seg_done = 0;
    auto seg_lo = [&](uint32_t sg) { return (uint32_t)((uint64_t)count * sg / nseg); };
    const uint32_t seg_waves = max(1u, waves / nseg);
    bool queue_open = true;
    uint32_t lq = 0, lq_end = 0, rem = seg_lo(seg + 1) - seg_lo(seg);
    // parked pieces of earlier rounds (taken after the queue's items)
    const uint32_t cont_count = a.cont_in ? *a.cont_in_count : 0u;
    bool cont_open = cont_count != 0;
    uint64_t pass = 0;  // wave-uniform: passes run so far

    // Next item (single-launch drains): a lane that holds a piece also holds
    // the queue item it starts when the piece ends. When that item has the
    // lane's solve and weight (bit set in hasnext), the search phase starts
    // it at once, so the lane does not idle until the phase ends; otherwise
    // the lane starts it at the next refill.
    const bool prefetch = a.budget == 0;
    uint4 nx = make_uint4(0u, 0u, 0u, 0u);  // {f, avail, ld, rd} of the next item
    uint32_t nsolve = 0, nws = 0, nmask = 0;
    bool holdnext = false;
    uint64_t hasnext = 0;  // wave-uniform: lanes whose next item the phase may start

    // resume parked pieces exactly where they stopped: current row in VGPRs,
    // open rows to LDS
    auto resume_parked = [&](uint64_t& idle) {
        const uint32_t nidle = __popcll(idle);
        uint32_t cb = 0;
        if (lane == 0) cb = atomicAdd(a.cont_head, nidle);
        cb = __shfl(cb, 0, 64);
        const uint32_t cgot = cb >= cont_count ? 0u : min(nidle, cont_count - cb);
        if (cgot < nidle) cont_open = false;
        const uint32_t r = lane_rank(idle);
        const bool take = !holds() && r < cgot;
        const Cont* const c = a.cont_in + cb + r;
        const uint32_t csolve = take ? c->solve : 0u;
        switch_solve(a, take && csolve != acc_solve, csolve, acc_solve, acc_sols, acc_nodes, lane);
        if (take) {
            const uint32_t cmeta = c->meta;
            uint32_t k = c->k;
            if (k > a.stack_rows) {  // cannot happen with a consistent stack depth
                atomicOr(a.err, 32u);
                k = 0;
            }
            mask = (1u << meta_n(cmeta)) - 1u;
            solve = csolve;
            ws = meta_wshift(cmeta);
            st = *reinterpret_cast<const uint4*>(c->st);
            for (uint32_t j = 0; j < k; ++j) slot_store(base + row * (j + 1u), *reinterpret_cast<const uint4*>(c->rows[j]));
            q = base + k * row;
        }
        idle = __ballot(!holds());
    };

    for (;;) {
        // ---- pass budget: park every unfinished piece, hand back every held
        // item, then leave
        if (a.budget && pass >= a.budget) {
            const bool park = holds();
            // once the queue and the parked pieces are used up, lanes would
            // idle next round: a piece with an open row parks its bottom row
            // (its largest subtree) as a piece of its own, so the next round
            // spreads the work again
            const bool split = park && !queue_open && lq == lq_end && !cont_open && q != base;
            const uint64_t pm = __ballot(park), sm = __ballot(split);
            const uint32_t np = (uint32_t)__popcll(pm) + (uint32_t)__popcll(sm);
            if (np) {
                uint32_t o = 0;
                if (lane == 0) o = atomicAdd(a.cont_out_count, np);
                o = __shfl(o, 0, 64);
                if ((uint64_t)o + np > a.cont_cap) {
                    if (lane == 0) atomicOr(a.err, 4u);
                } else if (park) {
                    const uint32_t meta = item_meta(31u - __clz(mask + 1u), 0u, ws);
                    uint32_t bot = base + row;  // slot 0
                    if (split) {
                        Cont* c = a.cont_out + o + (uint32_t)__popcll(pm) + lane_rank(sm);
                        c->solve = solve;
                        c->meta = meta;
                        c->k = 0;
                        c->pad = 0;
                        *reinterpret_cast<uint4*>(c->st) = slot_load(bot);
                        bot += row;
                    }
                    Cont* c = a.cont_out + o + lane_rank(pm);
                    const uint32_t k = (q + row - bot) / row;
                    c->solve = solve;
                    c->meta = meta;
                    c->k = k;
                    c->pad = 0;
                    *reinterpret_cast<uint4*>(c->st) = st;
                    for (uint32_t j = 0; j < k; ++j) *reinterpret_cast<uint4*>(c->rows[j]) = slot_load(bot + j * row);
                }
            }
            st.x = 0;
            q = base;
            if (lq < lq_end) {  // taken but never started: back to the queue
                const uint32_t nloc = lq_end - lq;
                uint32_t o = 0;
                if (lane == 0) o = atomicAdd(a.spill_count, nloc);
                o = __shfl(o, 0, 64);
                if ((uint64_t)o + nloc > a.spill_cap) {
                    if (lane == 0) atomicOr(a.err, 4u);
                } else {
                    for (uint32_t j = lane; j < nloc; j += 64u) a.spill[o + j] = a.items[lq + j];
                }
                lq = lq_end;
            }
            break;
        }

        // ---- refill idle lanes: parked pieces of earlier rounds, then held
        // items, then the global queue (one atomic per grab of up to kGrabMax
        // items), then in-wave donation

        // The queue items for the lanes of `want` (wave-collective): the
        // held range first, then a new grab; returns the lane's item index
        // or kNoItem.
        constexpr uint32_t kNoItem = 0xffffffffu;
        auto take_queue = [&](uint64_t want) -> uint32_t {
            const uint32_t nw = __popcll(want);
            const uint32_t take_local = min(nw, lq_end - lq);
            uint32_t g = 0, got = 0;
            if (take_local < nw && queue_open) {
                // guided self-scheduling: a share of what is left, at
                // least what the lanes need
                const uint32_t need = nw - take_local;
                uint32_t wantn = rem / (2u * seg_waves);
                wantn = wantn < need ? need : (wantn > kGrabMax ? kGrabMax : wantn);
                const uint32_t lo = seg_lo(seg), len = seg_lo(seg + 1) - lo;
                if (lane == 0) g = atomicAdd(a.head + seg * kHeadStride, wantn);
                g = __shfl(g, 0, 64);
                got = g >= len ? 0u : min(wantn, len - g);
                rem = g >= len ? 0u : len - g - got;
                g += lo;
                if (got < wantn) {  // this segment is used up: the next one
                    if (++seg_done == nseg) queue_open = false;
                    seg = seg + 1u == nseg ? 0u : seg + 1u;
                    rem = seg_lo(seg + 1) - seg_lo(seg);
                }
            }
            const uint32_t r = lane_rank(want);
            uint32_t idx = kNoItem;
            if ((want >> lane) & 1u) {
                if (r < take_local)
                    idx = lq + r;
                else if (r - take_local < got)
                    idx = g + (r - take_local);
            }
            lq += take_local;
            if (got) {  // the held range was used up: keep the rest of the grab
                lq = g + min(got, nw - take_local);
                lq_end = g + got;
            }
            return idx;
        };

        // a lane whose piece ended and whose next item the phase could not
        // start (another solve or weight) starts it now
        if (prefetch) {
            const bool use = !holds() && holdnext;
            switch_solve(a, use && nsolve != acc_solve, nsolve, acc_solve, acc_sols, acc_nodes, lane);
            if (use) {
                mask = nmask;
                st = nx;
                solve = nsolve;
                ws = nws;
                q = base;
                holdnext = false;
            }
        }
        uint64_t idle = __ballot(!holds());
        if (idle) {
            if (idle && cont_open && !a.items_first) resume_parked(idle);
            if (idle && (queue_open || lq < lq_end)) {
                const uint32_t idx = take_queue(idle);
                const bool take = idx != kNoItem;
                Item it{};
                if (take) it = a.items[idx];
                switch_solve(a, take && it.solve != acc_solve, it.solve, acc_solve, acc_sols, acc_nodes, lane);
                if (take) {
                    mask = (1u << meta_n(it.meta)) - 1u;
                    st = make_uint4(it.f, ~it.cols & mask, it.ld, it.rd);
                    solve = it.solve;
                    ws = meta_wshift(it.meta);
                    q = base;
                }
                idle = __ballot(!holds());
            }
            if (idle && cont_open && a.items_first) resume_parked(idle);
            if (idle && a.donate) {
                // a lane with an open row gives its bottom one (its largest
                // untried subtree) to an idle lane, then moves its other open
                // rows down one slot, so every stack still starts at its base
                const bool is_donor = q != base;
                const uint64_t donors = __ballot(is_donor);
                const uint32_t k = min((uint32_t)__popcll(idle), (uint32_t)__popcll(donors));
                if (k) {
                    const uint32_t dr = lane_rank(donors);
                    const bool gives = is_donor && dr < k;
                    if (gives) wscratch[dr] = lane;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    const uint32_t ir = lane_rank(idle);
                    const bool recv = !holds() && ir < k;
                    const uint32_t dl = recv ? wscratch[ir] : lane;
                    const uint32_t d_mask = __shfl(mask, dl, 64);
                    const uint32_t d_solve = __shfl(solve, dl, 64);
                    const uint32_t d_ws = __shfl(ws, dl, 64);
                    switch_solve(a, recv && d_solve != acc_solve, d_solve, acc_solve, acc_sols, acc_nodes, lane);
                    if (recv) {
                        // the donor's bottom open row, straight from its LDS slot
                        st = slot_load(base - 16u * lane + 16u * dl + row);  // the donor's slot 0
                        mask = d_mask;
                        solve = d_solve;
                        ws = d_ws;
                        q = base;
                    }
                    // every receiver has read its entry before the donors overwrite it
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    if (gives) {
                        for (uint32_t s = base + 2u * row; s <= q; s += row) slot_store(s - row, slot_load(s));
                        q -= row;
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
        }
        // lanes with a piece and no next item take one
        if (prefetch && (queue_open || lq < lq_end)) {
            const uint64_t want = __ballot(holds() && !holdnext);
            if (want) {
                const uint32_t idx = take_queue(want);
                if (idx != kNoItem) {
                    const Item it = a.items[idx];
                    nmask = (1u << meta_n(it.meta)) - 1u;
                    nx = make_uint4(it.f, ~it.cols & nmask, it.ld, it.rd);
                    nsolve = it.solve;
                    nws = meta_wshift(it.meta);
                    holdnext = true;
                }
            }
        }
        hasnext = __ballot(holdnext && nsolve == solve && nws == ws);

        const uint64_t live_mask = __ballot(holds());
        const bool have_items = queue_open || lq < lq_end || cont_open;
        if (live_mask == 0 && !have_items) break;

        // ---- search: every lane with a piece places one queen per pass (the
        // K1 step) until enough lanes idle to make a refill worth it (only
        // while a refill can find work), every lane is idle, or the phase
        // limit is reached
        const uint32_t nlive = __popcll(live_mask);
        const bool can_refill = have_items || (a.donate && __ballot(q != base) != 0);
        uint32_t floor_live = 0;
        uint64_t limit = a.budget ? a.budget - pass : (uint64_t)kPhaseCap;
        if (can_refill) {
            if (nlive > 64u - a.refill_idle)
                floor_live = 64u - a.refill_idle;  // leave once refill_idle lanes idle
            else
                limit = limit < 256u ? limit : 256u;  // idle lanes already: retry soon
        }
        floor_live = __builtin_amdgcn_readfirstlane(floor_live);
        const uint32_t lim32 = __builtin_amdgcn_readfirstlane(
            (uint32_t)(limit < (uint64_t)kPhaseCap ? limit : (uint64_t)kPhaseCap));
        // one (solve, weight) for every lane with a piece: the phase counts
        // per wave (lane masks, scalar sums), else per lane
        const uint32_t lead = live_mask ? (uint32_t)__ffsll((long long)live_mask) - 1u : 0u;
        const uint32_t lead_solve = __shfl(solve, lead, 64);
        const uint32_t lead_ws = __shfl(ws, lead, 64);
        const bool live = holds();
        const bool uniform = __ballot(live && (solve != lead_solve || ws != lead_ws)) == 0;
        uint32_t psols = 0, pnodes = 0, ran = 0;  // this phase (<= kPhaseCap passes: no wrap)
        // the phase also ends once next_refill lanes have started their next
        // item (while the queue can give them another)
        const int32_t nfloor = __builtin_amdgcn_readfirstlane(
            (queue_open || lq < lq_end) ? (int32_t)__popcll(hasnext) - (int32_t)a.next_refill : -1);
        if (live) {
            const uint64_t had = hasnext;
            if (uniform)
                search_phase<T, true>(st, q, nx, hasnext, nfloor, floor_live, lim32, psols, pnodes, ran);
            else
                search_phase<T, false>(st, q, nx, hasnext, nfloor, floor_live, lim32, psols, pnodes, ran);
            // the phase cleared the bits of the lanes that started their next item
            if (((had & ~hasnext) >> lane) & 1u) holdnext = false;
            // a lane whose piece ended borrowed below its stack; its last
            // load (its own row-0 slot) is void
            if ((int32_t)q < 0) {
                q = base;
                st.x = 0;
            }
        }
        ran = wave_max_u32(ran);  // lanes without a piece did not enter
        if (uniform) {  // the wave's counts, once
            psols = lane == lead ? psols : 0u;
            pnodes = lane == lead ? pnodes : 0u;
        }
        exec_nodes += pnodes;
        acc_nodes += (uint64_t)pnodes << ws;
        acc_sols += (uint64_t)psols << ws;
        pass += ran;
    }

    exec_nodes = wave_sum64(exec_nodes);
    if (a.wave_part) {
        // Store the wave's totals; k2_drain_commit adds them up after the
        // launch. Thousands of waves ending together would otherwise queue
        // on the same three counters.
        const uint64_t act = __ballot((acc_sols | acc_nodes) != 0);
        const uint32_t first = act ? (uint32_t)__ffsll((long long)act) - 1u : 0u;
        const uint32_t s0 = __shfl(acc_solve, first, 64);
        const bool one = __ballot((acc_sols | acc_nodes) != 0 && acc_solve != s0) == 0;
        uint64_t ws_sols = 0, ws_nodes = 0;
        if (one) {
            ws_sols = wave_sum64(acc_sols);
            ws_nodes = wave_sum64(acc_nodes);
        } else {  // several solves (rare): add them directly
            add_per_solve(a.solve_sols, acc_solve, acc_sols, lane);
            add_per_solve(a.solve_nodes, acc_solve, acc_nodes, lane);
        }
        if (lane == 0) {
            uint64_t* const r = a.wave_part + 4u * (blockIdx.x * (T / 64u) + wave);
            r[0] = s0;
            r[1] = ws_sols;
            r[2] = ws_nodes;
            r[3] = exec_nodes;
        }
    } else {
        add_per_solve(a.solve_sols, acc_solve, acc_sols, lane);
        add_per_solve(a.solve_nodes, acc_solve, acc_nodes, lane);
        if (lane == 0 && exec_nodes)
            atomicAdd((unsigned long long*)a.exec_nodes, (unsigned long long)exec_nodes);
    }
}

// One block: the per-wave totals of a drain launch into the per-solve
// counters. Neighbouring waves mostly share a solve, so each thread sums a
// run of waves per solve and the waves add those up (a few atomics).
__global__ void __launch_bounds__(256)
k2_drain_commit_kernel(const uint64_t* __restrict__ part, uint32_t nwaves, uint64_t* __restrict__ solve_sols,
                       uint64_t* __restrict__ solve_nodes, uint64_t* __restrict__ exec_nodes) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t per = (nwaves + 255u) / 256u;
    const uint32_t j0 = min(nwaves, threadIdx.x * per), j1 = min(nwaves, j0 + per);
    uint32_t acc_solve = 0;
    uint64_t acc_sols = 0, acc_nodes = 0, acc_exec = 0;
    for (uint32_t j = j0; j < j1; ++j) {
        const uint32_t s = (uint32_t)part[4u * j];
        const uint64_t ps = part[4u * j + 1u], pn = part[4u * j + 2u];
        acc_exec += part[4u * j + 3u];
        if ((ps | pn) == 0) continue;
        if (s != acc_solve && (acc_sols | acc_nodes) != 0) {
            if (acc_sols) atomicAdd((unsigned long long*)&solve_sols[acc_solve], (unsigned long long)acc_sols);
            if (acc_nodes) atomicAdd((unsigned long long*)&solve_nodes[acc_solve], (unsigned long long)acc_nodes);
            acc_sols = acc_nodes = 0;
        }
        acc_solve = s;
        acc_sols += ps;
        acc_nodes += pn;
    }
    add_per_solve(solve_sols, acc_solve, acc_sols, lane);
    add_per_solve(solve_nodes, acc_solve, acc_nodes, lane);
    const uint64_t wexec = wave_sum64(acc_exec);
    if (lane == 0 && wexec) atomicAdd((unsigned long long*)exec_nodes, (unsigned long long)wexec);
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

hipError_t k2_expand_count(hipStream_t s, uint32_t grid, const Item* in, const uint32_t* in_count, void* part,
                           uint32_t* out_count, const uint8_t* target_row, const float* pred,
                           const float* thresh) {
    if (grid == 0 || grid > kExpandMaxGrid) return hipErrorInvalidValue;
    ExpandPart* const p = static_cast<ExpandPart*>(part);
    hipLaunchKernelGGL(k2_expand_count_kernel, dim3(grid), dim3(256), 0, s, in, in_count, p, target_row, pred,
                       thresh);
    hipLaunchKernelGGL(k2_expand_total_kernel, dim3(1), dim3(256), 0, s, p, grid, out_count);
    return hipGetLastError();
}

hipError_t k2_expand_write(hipStream_t s, uint32_t grid, const Item* in, const uint32_t* in_count,
                           const void* part, bool interleave, Item* out, uint32_t out_cap,
                           const uint8_t* target_row, uint64_t* solve_nodes, uint64_t* exec_nodes, uint32_t* err,
                           const float* pred, const float* thresh) {
    if (grid == 0 || grid > kExpandMaxGrid) return hipErrorInvalidValue;
    const ExpandPart* const p = static_cast<const ExpandPart*>(part);
    hipLaunchKernelGGL(k2_expand_write_kernel, dim3(grid), dim3(256), 0, s, in, in_count, p,
                       interleave ? 1u : 0u, out, out_cap, target_row, solve_nodes, exec_nodes, err, pred, thresh);
    hipLaunchKernelGGL(k2_expand_commit_kernel, dim3(1), dim3(256), 0, s, p, grid, solve_nodes, exec_nodes);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// largest-predicted-first order: counting sort on 64 buckets of half a log2
// ---------------------------------------------------------------------------
constexpr uint32_t kBuckets = 64;

__device__ __forceinline__ uint32_t pred_bucket(float p) {
    // bucket 0 holds the largest predictions (log2 sizes 31.5 and up)
    // (a NaN prediction goes to bucket 0)
    const float b = (32.0f - p) * 2.0f;
    return !(b > 0.f) ? 0u : (b >= (float)(kBuckets - 1) ? kBuckets - 1 : (uint32_t)b);
}

__global__ void k2_bucket_hist_kernel(const float* __restrict__ pred, const uint32_t* __restrict__ count,
                                      uint32_t max_count, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kBuckets];
    for (uint32_t i = threadIdx.x; i < kBuckets; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const uint32_t c = min(*count, max_count);
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

// One global atomic per (block, bucket): the block counts its items per
// bucket in LDS, reserves each bucket's range once, then ranks its items
// inside the range with LDS atomics. Order inside a bucket is arbitrary.
// The ranks are dealt round-robin to nseg equal runs (rank r goes to run
// r % nseg), so that each queue segment of the drain holds the same mix of
// sizes, largest first, instead of one segment holding all the largest.
__global__ void __launch_bounds__(256)
k2_bucket_scatter_kernel(const Item* __restrict__ in, const float* __restrict__ pred,
                         const uint32_t* __restrict__ count, uint32_t max_count, uint32_t nseg,
                         uint32_t* __restrict__ cursor, Item* __restrict__ out) {
    __shared__ uint32_t h[kBuckets];
    __shared__ uint32_t off[kBuckets];
    const uint32_t c = min(*count, max_count);
    // run s holds the ranks r = s (mod nseg): q + (s < rm) of them, from s q + min(s, rm)
    const uint32_t q = c / nseg, rm = c % nseg;
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < c; b0 += gridDim.x * blockDim.x) {
        for (uint32_t i = threadIdx.x; i < kBuckets; i += blockDim.x) h[i] = 0;
        __syncthreads();
        const uint32_t i = b0 + threadIdx.x;
        uint32_t bk = 0, r = 0;
        if (i < c) {
            bk = pred_bucket(pred[i]);
            r = atomicAdd(&h[bk], 1u);
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < kBuckets; j += blockDim.x)
            off[j] = h[j] ? atomicAdd(&cursor[j], h[j]) : 0u;
        __syncthreads();
        if (i < c) {
            const uint32_t rank = off[bk] + r, s = rank % nseg;
            out[s * q + min(s, rm) + rank / nseg] = in[i];
        }
        __syncthreads();
    }
}

hipError_t k2_order_by_prediction(hipStream_t s, const Item* in, const float* pred, const uint32_t* count,
                                  uint32_t max_count, uint32_t nseg, uint32_t* hist, Item* out) {
    if (nseg == 0) return hipErrorInvalidValue;
    const uint32_t blocks = std::max(1u, std::min(2048u, (max_count + 255u) / 256u));
    hipLaunchKernelGGL(k2_fill_u32_kernel, dim3(1), dim3(64), 0, s, hist, kBuckets, 0u);
    hipLaunchKernelGGL(k2_bucket_hist_kernel, dim3(blocks), dim3(256), 0, s, pred, count, max_count, hist);
    hipLaunchKernelGGL(k2_bucket_scan_kernel, dim3(1), dim3(64), 0, s, hist);
    hipLaunchKernelGGL(k2_bucket_scatter_kernel, dim3(blocks), dim3(256), 0, s, in, pred, count, max_count, nseg,
                       hist, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// learned split threshold, computed where the predictions are
// ---------------------------------------------------------------------------
// The sum of 2^pred is formed in a fixed order (per-block partials, then
// one block adds them), so that the threshold, and with it every size the
// split decides, is the same on every run over the same items: a captured
// graph replays the sizes of the run that planned it.
__device__ __forceinline__ double block_sum256_f64(double v, double* part) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (uint32_t w = 0; w < blockDim.x / 64u; ++w) t += part[w];
    return t;
}

__global__ void __launch_bounds__(256)
k2_pred_sum_kernel(const float* __restrict__ pred, const uint32_t* __restrict__ count, uint32_t max_count,
                   double* __restrict__ block_sums) {
    __shared__ double part[4];
    const uint32_t c = min(*count, max_count);
    double v = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < c; i += gridDim.x * blockDim.x)
        v += exp2((double)pred[i]);
    const double t = block_sum256_f64(v, part);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = t;
}

__global__ void __launch_bounds__(256)
k2_pred_total_kernel(const double* __restrict__ block_sums, uint32_t blocks, double* __restrict__ sum) {
    __shared__ double part[4];
    double v = 0.0;
    for (uint32_t i = threadIdx.x; i < blocks; i += blockDim.x) v += block_sums[i];
    const double t = block_sum256_f64(v, part);
    if (threadIdx.x == 0) *sum = t;
}

__device__ __forceinline__ float split_thresh(float fixed, double mult, const double* sum, uint32_t count) {
    return fixed > 0.f ? fixed : (float)log2(fmax(1.0, *sum * mult / (double)max(count, 1u)));
}

__global__ void __launch_bounds__(256)
k2_pred_over_kernel(const float* __restrict__ pred, const uint32_t* __restrict__ count, uint32_t max_count,
                    float fixed, double mult, const double* __restrict__ sum, uint32_t* __restrict__ over,
                    float* __restrict__ thresh) {
    const uint32_t c = min(*count, max_count);
    const float t = split_thresh(fixed, mult, sum, c);
    if (blockIdx.x == 0 && threadIdx.x == 0) *thresh = t;
    uint32_t n = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < c; i += gridDim.x * blockDim.x)
        n += pred[i] > t;
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o, 64);
    if ((threadIdx.x & 63u) == 0 && n) atomicAdd(over, n);
}

hipError_t k2_split_threshold(hipStream_t s, const float* pred, const uint32_t* count, uint32_t max_count,
                              float fixed, double mult, double* block_sums, double* sum, uint32_t* over,
                              float* thresh) {
    const uint32_t blocks = std::max(1u, std::min(kPredSumBlocks, (max_count + 255u) / 256u));
    if (!(fixed > 0.f)) {
        hipLaunchKernelGGL(k2_pred_sum_kernel, dim3(blocks), dim3(256), 0, s, pred, count, max_count, block_sums);
        hipLaunchKernelGGL(k2_pred_total_kernel, dim3(1), dim3(256), 0, s, block_sums, blocks, sum);
    }
    hipLaunchKernelGGL(k2_pred_over_kernel, dim3(blocks), dim3(256), 0, s, pred, count, max_count, fixed, mult,
                       sum, over, thresh);
    return hipGetLastError();
}

size_t k2_lds_bytes(uint32_t stack_rows, uint32_t threads) {
    return threads == 64 ? Stack<64>::bytes(stack_rows)
                         : threads == 128 ? Stack<128>::bytes(stack_rows) : Stack<256>::bytes(stack_rows);
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
    a.next_refill = p.next_refill ? p.next_refill : std::min(56u, 3u * p.refill_idle);
    a.donate = p.donate;
    a.budget = p.budget;
    a.cont_in = p.cont_in;
    a.cont_in_count = p.cont_in_count;
    a.cont_head = p.cont_head;
    a.cont_out = p.cont_out;
    a.cont_out_count = p.cont_out_count;
    a.cont_cap = p.cont_cap;
    a.items_first = p.items_first;
    a.nseg = p.nseg ? p.nseg : 1u;
    a.wave_part = p.wave_part;
    const size_t lds = k2_lds_bytes(p.stack_rows, threads);
    if (threads == 64)
        hipLaunchKernelGGL(k2_drain_kernel<64>, dim3(blocks), dim3(64), lds, s, a);
    else if (threads == 128)
        hipLaunchKernelGGL(k2_drain_kernel<128>, dim3(blocks), dim3(128), lds, s, a);
    else if (threads == 256)
        hipLaunchKernelGGL(k2_drain_kernel<256>, dim3(blocks), dim3(256), lds, s, a);
    else
        return hipErrorInvalidValue;
    if (p.wave_part)
        hipLaunchKernelGGL(k2_drain_commit_kernel, dim3(1), dim3(256), 0, s, p.wave_part,
                           blocks * (threads / 64u), p.solve_sols, p.solve_nodes, p.exec_nodes);
    return hipGetLastError();
}

int k2_auto_threads(uint32_t stack_rows, uint32_t lds_per_block, uint32_t* threads) {
    // resident workgroups per CU for workgroup size c (0: the stack does not fit)
    auto blocks_per_cu = [&](uint32_t c, int* nb) -> int {
        *nb = 0;
        if (k2_lds_bytes(stack_rows, c) > lds_per_block) return NQM_OK;
        return k2_max_blocks_per_cu(c, k2_lds_bytes(stack_rows, c), nb);
    };
    int rc, nb64 = 0;
    if ((rc = blocks_per_cu(64, &nb64)) != NQM_OK) return rc;
    *threads = 64;
    for (uint32_t c : {128u, 256u}) {
        int nb = 0;
        if ((rc = blocks_per_cu(c, &nb)) != NQM_OK) return rc;
        if (nb > 0 && nb * (int)(c / 64u) >= nb64 - 1) *threads = c;
    }
    return NQM_OK;
}

int k2_max_blocks_per_cu(uint32_t threads, size_t lds_bytes, int* out) {
    int nb = 0;
    hipError_t e = threads == 64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k2_drain_kernel<64>, 64, lds_bytes)
                   : threads == 128
                       ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k2_drain_kernel<128>, 128, lds_bytes)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k2_drain_kernel<256>, 256, lds_bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return NQM_EDEVICE;
    }
    *out = nb;
    return NQM_OK;
}

}  // namespace nqm
