// k1_static.hip -- K1: static-pool counting kernel for gfx950 (wave64).
//
// One prefix per lane, as kern_doitall_v2_regld (GPU_NQueens.cu:338-452):
// the lane backtracks rows >= depth with unit-row propagation, the wave stops
// on a 64-lane ballot done-vote, counts are u64, one partial per block.
//
// MI355X design (not a translation of the CUDA kernel):
//  * Relative u32 bitsets live in VGPRs: avail (free columns, the complement
//    of the reference's queens_in_columns), ld, rd (columns attacked at the
//    current row along the two diagonals) and f, the candidates of the
//    current row not tried yet. The free set of a row is one v_bitop3:
//    avail & ~ld & ~rd.
//  * One pass = one placement per live lane: take b = f & -f and compute the
//    child (avail^b, (ld|b)<<1, (rd|b)>>1) and its free set f'. A child with
//    f' == 0 is a leaf: it is counted (a solution when avail^b == 0, i.e. all
//    columns are taken) and the lane stays on its row. Only a child with a
//    free cell is entered. A unit row (one candidate) is entered without
//    saving anything: this is the reference's DoubleSweep-light propagation
//    (GPU_NQueens.cu:392-408); the search tree, hence the counts and node
//    numbers, are those of plain backtracking.
//  * The backtrack stack is in LDS, one 16-byte entry {f, avail, ld, rd} per
//    open row that still has untried candidates, laid out [row][lane] so a
//    wave's ds_write_b128 / ds_read_b128 touches 1 KiB of consecutive banks.
//    Leaves and unit rows never touch LDS; a lane pops only when its row is
//    exhausted, and the popped row always has a candidate left.
//  * Nodes are not counted per placement: a lane is live from pass 0 and
//    places exactly one queen per pass, so its node count is the wave-uniform
//    (scalar) pass counter at the moment it finishes.
//  * Wave sum of u64 through cross-lane shuffles, LDS across the block, one
//    16-byte store per block (the reference does 1024 same-address atomics,
//    GPU_NQueens.cu:441-443). Workgroups go round-robin to the 8 XCDs, each
//    with its own L2, so the records are grouped per XCD (k1_part_slot): a
//    cache line of records is written by one XCD only, not by eight.

#include <type_traits>

#include "nqm_internal.h"

namespace nqm {

constexpr uint32_t kK1MaxThreads = 256;

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum v over the block; the result is valid in thread 0. `scratch` needs
// 2 * (blockDim.x / 64) u64 of LDS and must not be in use by other lanes.
__device__ __forceinline__ void block_sum2_u64(uint64_t& a, uint64_t& b, uint64_t* scratch) {
    a = wave_sum_u64(a);
    b = wave_sum_u64(b);
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t nwaves = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) {
        scratch[wave] = a;
        scratch[nwaves + wave] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < nwaves; ++w) {
            a += scratch[w];
            b += scratch[nwaves + w];
        }
    }
}

// kWide: u64 solution counter per lane. Without it the counter is u32, which
// the host allows only when the whole board has fewer than 2^32 solutions
// (n <= 18, test.cu:18), so no prefix can overflow it.
template <bool kWide>
__global__ void __launch_bounds__(kK1MaxThreads)
k1_count_kernel(const uint8_t* __restrict__ pool, uint64_t count, uint32_t n, uint32_t depth,
                uint32_t flags, uint64_t* __restrict__ block_part, uint64_t* __restrict__ per_sols,
                uint64_t* __restrict__ per_nodes, uint32_t* __restrict__ err) {
    extern __shared__ uint4 lds_stack[];  // [rows][blockDim.x]
    const uint32_t tid = threadIdx.x;
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + tid;
    const uint32_t mask = (1u << n) - 1u;  // n <= 31

    // st = {f, avail, ld, rd}: the current row's untried candidates, the free
    // columns and the two diagonal attack sets at that row
    uint4 st = make_uint4(0, 0, 0, 0);
    uint32_t c0 = 0;
    bool live = false;
    if (gid < count) {
        const uint8_t* p = pool + gid * depth;
        uint32_t cols = 0, ld = 0, rd = 0;
        bool ok = true;
        for (uint32_t r = 0; r < depth; ++r) {
            const uint32_t c = p[r];
            if (r == 0) c0 = c;
            const uint32_t b = 1u << (c & 31u);
            ok = ok && (c < n) && !(b & (cols | ld | rd));
            cols |= b;
            ld = (ld | b) << 1;
            rd = (rd | b) >> 1;
        }
        if (!ok) {
            atomicOr(err, 1u);
        } else {
            live = true;
            const uint32_t avail = ~cols & mask;
            st = make_uint4(avail & ~ld & ~rd, avail, ld, rd);
        }
    }

    // LDS byte address of this lane's next free stack slot; slots of one lane
    // are blockDim.x * 16 bytes apart.
    const uint32_t step = blockDim.x * (uint32_t)sizeof(uint4);
    const uint32_t base = tid * (uint32_t)sizeof(uint4);
    char* const lds = reinterpret_cast<char*>(lds_stack);
    uint64_t sols = 0, nodes = 0;
    typename std::conditional<kWide, uint64_t, uint32_t>::type lsols = 0;

    // The per-lane loop runs while the lane has work; the wave leaves it when
    // its execution mask is empty, which is the 64-lane done-vote of the
    // reference (GPU_NQueens.cu:359-360).
    if (live) {
        // a complete prefix (depth == n) is one solution with nothing below it
        lsols = st.y == 0 ? 1u : 0u;
        uint32_t top = base;
        for (;;) {
            if (st.x == 0) {  // row exhausted: resume the nearest open row
                if (top == base) break;
                top -= step;
                st = *reinterpret_cast<const uint4*>(lds + top);
            }
            const uint32_t b = st.x & (0u - st.x);
            st.x ^= b;
            const uint32_t a2 = st.y ^ b;
            const uint32_t l2 = (st.z + b) << 1;  // b is not in ld: + is |
            const uint32_t r2 = (st.w | b) >> 1;
            const uint32_t f2 = a2 & ~l2 & ~r2;
            lsols += (a2 == 0);  // all columns taken: a solution (a leaf)
            const bool go = f2 != 0;
            if (go && st.x != 0) {  // entering the child: save this row, it has more
                *reinterpret_cast<uint4*>(lds + top) = st;
                top += step;
            }
            st.x = go ? f2 : st.x;
            st.y = go ? a2 : st.y;
            st.z = go ? l2 : st.z;
            st.w = go ? r2 : st.w;
            ++nodes;  // one placement per pass
        }
        sols = lsols;
    }

    if (gid < count) {
        if (per_sols) per_sols[gid] = sols;
        if (per_nodes) per_nodes[gid] = nodes;
    }
    // GPU_NQueens.cu:356, :437-439: x2 for row-0 columns left of n/2.
    if ((flags & NQM_MIRROR_WEIGHT) && depth > 0 && c0 < (n >> 1)) sols <<= 1;

    __syncthreads();  // LDS stack is reused as reduction scratch
    block_sum2_u64(sols, nodes, reinterpret_cast<uint64_t*>(lds_stack));
    if (tid == 0) {
        uint64_t* const r = block_part + 2u * k1_part_slot(blockIdx.x, gridDim.x);
        r[0] = sols;
        r[1] = nodes;
    }
}

int k1_plan(const nqm_ctx* ctx, uint32_t n, uint32_t depth, uint64_t count, K1Launch* plan) {
    // Entries are pushed at rows depth..n-2 only (row n-1 has <= 1 candidate).
    const uint32_t rows = (n > depth + 1) ? (n - 1 - depth) : 1;
    // one-wave workgroups pack the per-lane LDS stacks of a CU most tightly
    uint32_t threads = 64;
    const size_t lds = (size_t)rows * threads * sizeof(uint4);
    if (lds > ctx->lds_per_block) return NQM_EUNSUPPORTED;
    const uint64_t blocks = (count + threads - 1) / threads;
    if (blocks > 0x7fffffffull) return NQM_ERANGE;
    plan->blocks = (uint32_t)blocks;
    plan->threads = threads;
    plan->lds_bytes = lds < 2 * (threads / 64) * sizeof(uint64_t) ? 2 * (threads / 64) * sizeof(uint64_t) : lds;
    return NQM_OK;
}

hipError_t k1_launch(const K1Launch& plan, hipStream_t stream, const uint8_t* d_pool,
                     uint64_t count, uint32_t n, uint32_t depth, uint32_t flags,
                     uint64_t* d_block_part, uint64_t* d_per_sols, uint64_t* d_per_nodes, uint32_t* d_err) {
    if (n <= 18)
        hipLaunchKernelGGL(k1_count_kernel<false>, dim3(plan.blocks), dim3(plan.threads), plan.lds_bytes,
                           stream, d_pool, count, n, depth, flags, d_block_part, d_per_sols, d_per_nodes, d_err);
    else
        hipLaunchKernelGGL(k1_count_kernel<true>, dim3(plan.blocks), dim3(plan.threads), plan.lds_bytes,
                           stream, d_pool, count, n, depth, flags, d_block_part, d_per_sols, d_per_nodes, d_err);
    return hipGetLastError();
}

}  // namespace nqm
