// k1_static.hip -- K1: static-pool counting kernel for gfx950 (wave64).
//
// One prefix per lane, as kern_doitall_v2_regld (GPU_NQueens.cu:338-452):
// the lane backtracks rows >= depth with unit-row propagation, the wave stops
// on a 64-lane ballot done-vote, counts are u64, one partial per block.
//
// MI355X design (not a translation of the CUDA kernel):
//  * Relative u32 bitsets (cols, ld, rd) live in VGPRs; the candidate mask f
//    of the current row is a fourth VGPR. A placement is
//      b = f & -f; f ^= b; cols |= b; ld = (ld|b)<<1; rd = (rd|b)>>1;
//      f' = ~(cols|ld|rd) & mask.
//  * The backtrack stack is in LDS, one 16-byte entry {f_rest, cols, ld, rd}
//    per open row, laid out [row][lane] so a wave's ds_write_b128 /
//    ds_read_b128 touches 1 KiB of consecutive banks. An entry is pushed only
//    when the row still has untried candidates, so a unit row (exactly one
//    free cell) is placed and passed without touching LDS: that is the
//    reference's DoubleSweep-light propagation (GPU_NQueens.cu:392-408), and
//    it leaves the search tree, hence counts and node numbers, unchanged.
//  * Every loop pass is one placement per live lane (pop first if the lane's
//    row is exhausted). Stack entries always hold a non-empty f, so one pop
//    suffices: no inner backtrack loop, no divergent while.
//  * Nodes are not counted per placement: a lane is live from pass 0 and
//    places exactly one queen per pass, so its node count is the wave-uniform
//    (scalar) pass counter at the moment it finishes.
//  * Solutions are detected when a lane's candidate mask is empty: the board
//    is full exactly when cols == mask (GPU_NQueens.cu:410-435).
//  * Wave sum of u64 through cross-lane shuffles, LDS across the block, one
//    store per block (the reference does 1024 same-address atomics,
//    GPU_NQueens.cu:441-443).

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

__global__ void __launch_bounds__(kK1MaxThreads)
k1_count_kernel(const uint8_t* __restrict__ pool, uint64_t count, uint32_t n, uint32_t depth,
                uint32_t flags, uint64_t* __restrict__ block_sols,
                uint64_t* __restrict__ block_nodes, uint64_t* __restrict__ per_sols,
                uint64_t* __restrict__ per_nodes, uint32_t* __restrict__ err) {
    extern __shared__ uint4 lds_stack[];  // [rows][blockDim.x]
    const uint32_t tid = threadIdx.x;
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + tid;
    const uint32_t mask = (1u << n) - 1u;  // n <= 31

    uint32_t cols = 0, ld = 0, rd = 0, f = 0;
    uint32_t c0 = 0;
    bool live = false;
    if (gid < count) {
        const uint8_t* p = pool + gid * depth;
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
            f = ~(cols | ld | rd) & mask;
        }
    }

    uint4* const my = lds_stack + tid;  // entry k of this lane: my[k * stride]
    const uint32_t stride = blockDim.x;
    uint32_t sp = 0;
    uint64_t sols = 0, nodes = 0;
    uint64_t pass = 0;  // wave-uniform

    for (;;) {
        if (live && f == 0) {
            sols += (cols == mask);
            if (sp == 0) {
                live = false;
                nodes = pass;
            } else {
                --sp;
                const uint4 e = my[sp * stride];
                f = e.x;
                cols = e.y;
                ld = e.z;
                rd = e.w;
            }
        }
        if (__ballot(live) == 0) break;  // 64-lane done-vote
        if (live) {
            const uint32_t b = f & (0u - f);
            f ^= b;
            if (f) {
                my[sp * stride] = make_uint4(f, cols, ld, rd);
                ++sp;
            }
            cols |= b;
            ld = (ld | b) << 1;
            rd = (rd | b) >> 1;
            f = ~(cols | ld | rd) & mask;
        }
        ++pass;
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
        block_sols[blockIdx.x] = sols;
        block_nodes[blockIdx.x] = nodes;
    }
}

int k1_plan(const nqm_ctx* ctx, uint32_t n, uint32_t depth, uint64_t count, K1Launch* plan) {
    // Entries are pushed at rows depth..n-2 only (row n-1 has <= 1 candidate).
    const uint32_t rows = (n > depth + 1) ? (n - 1 - depth) : 1;
    uint32_t threads = kK1MaxThreads;
    while (threads > 64 && (size_t)rows * threads * sizeof(uint4) > ctx->lds_per_block) threads >>= 1;
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
                     uint64_t* d_block_sols, uint64_t* d_block_nodes, uint64_t* d_per_sols,
                     uint64_t* d_per_nodes, uint32_t* d_err) {
    hipLaunchKernelGGL(k1_count_kernel, dim3(plan.blocks), dim3(plan.threads), plan.lds_bytes,
                       stream, d_pool, count, n, depth, flags, d_block_sols, d_block_nodes,
                       d_per_sols, d_per_nodes, d_err);
    return hipGetLastError();
}

}  // namespace nqm
