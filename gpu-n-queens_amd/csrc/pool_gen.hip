// pool_gen.hip -- prefix-pool generation on the device (reference order).
//
// Builds the pool the reference's host generator builds
// (gen_states_rowlock, n_queens.cu:27-83, one row-0 column per thread and
// concatenated in column order, n_queens.cu:111-124 and :203-213): every
// valid placement of queens in rows 0..lock-1 whose row `lock` still has a
// free cell, in lexicographic order of (col[0], col[1], ...). With mirror
// halving only row-0 columns < ceil(n/2) are used (n_queens.cu:130-134).
//
// MI355X design: a level-synchronous breadth-first expansion instead of one
// host thread per row-0 column. Each level runs three kernels: child counts
// per item with a block sum, one workgroup scanning the block sums, then the
// ordered write (block-local exclusive scan in LDS + the block's offset).
// The order of children (parent order, then column order) is exactly the
// reference's lexicographic order. A level item is 48 bytes: relative
// bitsets (cols, ld, rd) and the prefix's column bytes so far.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "nqm_internal.h"

namespace nqm {

namespace {

constexpr uint32_t kGenThreads = 256;

struct alignas(16) GenItem {
    uint32_t cols, ld, rd, row0;  // row0: column of the row-0 queen
    uint8_t q[32];                // columns of rows 0..row-1
};
static_assert(sizeof(GenItem) == 48, "GenItem layout");

__device__ __forceinline__ uint32_t child_mask(const GenItem& it, uint32_t mask) {
    // children whose next row keeps a free cell
    uint32_t f = ~(it.cols | it.ld | it.rd) & mask;
    uint32_t keep = 0;
    uint32_t g = f;
    while (g) {
        const uint32_t b = g & (0u - g);
        g ^= b;
        const uint32_t cf = ~((it.cols | b) | ((it.ld | b) << 1) | ((it.rd | b) >> 1)) & mask;
        if (cf) keep |= b;
    }
    return keep;
}

// Block-wide exclusive scan of one u32 per thread (256 threads), returns the
// exclusive prefix; *total receives the block sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) sh[wave] = x;
    __syncthreads();
    uint32_t wave_off = 0, tot = 0;
    for (uint32_t w = 0; w < blockDim.x / 64u; ++w) {
        if (w < wave) wave_off += sh[w];
        tot += sh[w];
    }
    __syncthreads();
    *total = tot;
    return wave_off + x - v;
}

__global__ void __launch_bounds__(kGenThreads)
gen_count_kernel(const GenItem* __restrict__ in, uint32_t count, uint32_t mask,
                 uint32_t* __restrict__ block_sums, unsigned long long* __restrict__ per_col) {
    __shared__ uint32_t sh[8];
    __shared__ unsigned long long cols[32];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (per_col && threadIdx.x < 32) cols[threadIdx.x] = 0;
    __syncthreads();
    uint32_t c = 0;
    if (i < count) {
        const GenItem it = in[i];
        c = __popc(child_mask(it, mask));
        // per row-0 column counts: LDS first, one global atomic per column
        // per block (items are grouped by column, so few columns per block)
        if (per_col && c) atomicAdd(&cols[it.row0], (unsigned long long)c);
    }
    uint32_t tot;
    (void)block_excl_scan(c, sh, &tot);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
    if (per_col && threadIdx.x < 32 && cols[threadIdx.x])
        atomicAdd(&per_col[threadIdx.x], cols[threadIdx.x]);
}

// One workgroup: exclusive scan of nblocks sums in place; writes the total.
__global__ void __launch_bounds__(1024)
gen_scan_kernel(uint32_t* __restrict__ sums, uint32_t nblocks, unsigned long long* __restrict__ total) {
    __shared__ unsigned long long sh[1024];
    __shared__ unsigned long long carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nblocks; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const unsigned long long v = i < nblocks ? sums[i] : 0ull;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t o = 1; o < 1024; o <<= 1) {
            const unsigned long long y = threadIdx.x >= o ? sh[threadIdx.x - o] : 0ull;
            __syncthreads();
            sh[threadIdx.x] += y;
            __syncthreads();
        }
        const unsigned long long incl = sh[threadIdx.x] + carry;
        if (i < nblocks) sums[i] = (uint32_t)(incl - v);  // offsets of this level fit u32
        __syncthreads();
        if (threadIdx.x == 1023) carry = incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

// Ordered write of children: to the next level (GenItem) or, on the last
// level, to the compact pool (depth bytes per prefix).
__global__ void __launch_bounds__(kGenThreads)
gen_write_kernel(const GenItem* __restrict__ in, uint32_t count, uint32_t mask, uint32_t row,
                 const uint32_t* __restrict__ block_off, GenItem* __restrict__ out_items,
                 uint8_t* __restrict__ out_pool, uint32_t depth) {
    __shared__ uint32_t sh[8];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    GenItem it;
    uint32_t keep = 0;
    if (i < count) {
        it = in[i];
        keep = child_mask(it, mask);
    }
    uint32_t tot;
    uint64_t o = (uint64_t)block_off[blockIdx.x] + block_excl_scan(__popc(keep), sh, &tot);
    while (keep) {
        const uint32_t b = keep & (0u - keep);
        keep ^= b;
        const uint32_t c = 31u - __clz(b);
        if (out_pool) {
            uint8_t* p = out_pool + o * depth;
            for (uint32_t r = 0; r < row; ++r) p[r] = it.q[r];
            p[row] = (uint8_t)c;
        } else {
            GenItem ch = it;
            ch.cols = it.cols | b;
            ch.ld = ((it.ld | b) << 1) & mask;
            ch.rd = (it.rd | b) >> 1;
            ch.q[row] = (uint8_t)c;
            out_items[o] = ch;
        }
        ++o;
    }
}

// Device buffers of one generation call.
struct GenBufs {
    DevBuf lvl[2], sums, total, per_col;
    ~GenBufs() {
        for (DevBuf* d : {&lvl[0], &lvl[1], &sums, &total, &per_col})
            if (d->ptr) (void)hipFree(d->ptr);
    }
};

// Runs the expansion down to `lock`. If out_pool is null only the count
// (and per-column counts) of the final level is produced.
int generate(nqm_ctx* ctx, uint32_t n, uint32_t lock, bool half, uint8_t* d_out_pool,
             uint64_t out_cap, uint64_t* count_out, uint64_t* per_col_out) {
    if (n < 2 || n > kMaxN || lock < 1 || lock >= n)
        return fail(ctx, NQM_EINVAL, "pool needs 2 <= n <= %u and 1 <= lock < n (n=%u lock=%u)", kMaxN,
                    n, lock);
    const uint32_t mask = (1u << n) - 1u;
    hipStream_t st = ctx->stream;
    GenBufs g;
    int rc;
    // level 1 items: one per row-0 column, built on the host
    const uint32_t span = half ? (n + 1) / 2 : n;
    std::vector<GenItem> l1(span);
    for (uint32_t c = 0; c < span; ++c) {
        GenItem it{};
        const uint32_t b = 1u << c;
        it.cols = b;
        it.ld = (b << 1) & mask;
        it.rd = b >> 1;
        it.row0 = c;
        it.q[0] = (uint8_t)c;
        l1[c] = it;
    }
    if ((rc = ensure_dev(ctx, g.lvl[0], span * sizeof(GenItem))) != NQM_OK) return rc;
    if ((rc = ensure_dev(ctx, g.total, 8)) != NQM_OK) return rc;
    if ((rc = ensure_dev(ctx, g.per_col, 32 * 8)) != NQM_OK) return rc;
    NQM_HIP(ctx, hipMemcpyAsync(g.lvl[0].ptr, l1.data(), span * sizeof(GenItem), hipMemcpyHostToDevice, st));
    NQM_HIP(ctx, hipMemsetAsync(g.per_col.ptr, 0, 32 * 8, st));

    // lock == 1: the pool is the row-0 queens whose row 1 keeps a free cell
    if (lock == 1) {
        uint64_t cnt = 0;
        std::vector<uint64_t> pc(n, 0);
        std::vector<uint8_t> host;
        for (uint32_t c = 0; c < span; ++c) {
            const GenItem& it = l1[c];
            if (~(it.cols | it.ld | it.rd) & mask) {
                pc[c] = 1;
                host.push_back((uint8_t)c);
                ++cnt;
            }
        }
        if (d_out_pool) {
            if (cnt > out_cap) return fail(ctx, NQM_ERANGE, "pool of %llu prefixes exceeds capacity",
                                           (unsigned long long)cnt);
            if (cnt) NQM_HIP(ctx, hipMemcpyAsync(d_out_pool, host.data(), cnt, hipMemcpyHostToDevice, st));
            NQM_HIP(ctx, hipStreamSynchronize(st));
        }
        *count_out = cnt;
        if (per_col_out) std::memcpy(per_col_out, pc.data(), n * sizeof(uint64_t));
        return NQM_OK;
    }

    uint64_t count = span;
    int cur = 0;
    for (uint32_t row = 1; row < lock; ++row) {
        // expanding level `row` items (rows 0..row-1 placed) into row+1
        const bool last = (row + 1 == lock);
        const uint64_t nblocks64 = (count + kGenThreads - 1) / kGenThreads;
        if (nblocks64 == 0) break;
        if (nblocks64 > 0x7fffffffull) return fail(ctx, NQM_ERANGE, "level too large");
        const uint32_t nblocks = (uint32_t)nblocks64;
        if ((rc = ensure_dev(ctx, g.sums, (size_t)nblocks * 4)) != NQM_OK) return rc;
        const GenItem* in = static_cast<const GenItem*>(g.lvl[cur].ptr);
        unsigned long long* pc = last ? static_cast<unsigned long long*>(g.per_col.ptr) : nullptr;
        hipLaunchKernelGGL(gen_count_kernel, dim3(nblocks), dim3(kGenThreads), 0, st, in, (uint32_t)count,
                           mask, static_cast<uint32_t*>(g.sums.ptr), pc);
        NQM_HIP(ctx, hipGetLastError());
        hipLaunchKernelGGL(gen_scan_kernel, dim3(1), dim3(1024), 0, st, static_cast<uint32_t*>(g.sums.ptr),
                           nblocks, static_cast<unsigned long long*>(g.total.ptr));
        NQM_HIP(ctx, hipGetLastError());
        uint64_t next = 0;
        NQM_HIP(ctx, hipMemcpyAsync(&next, g.total.ptr, 8, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipStreamSynchronize(st));
        if (next > 0xffffffffull) return fail(ctx, NQM_ERANGE, "level of %llu prefixes exceeds 2^32",
                                              (unsigned long long)next);
        if (last) {
            if (d_out_pool) {
                if (next > out_cap)
                    return fail(ctx, NQM_ERANGE, "pool of %llu prefixes exceeds capacity",
                                (unsigned long long)next);
                hipLaunchKernelGGL(gen_write_kernel, dim3(nblocks), dim3(kGenThreads), 0, st, in,
                                   (uint32_t)count, mask, row, static_cast<uint32_t*>(g.sums.ptr),
                                   (GenItem*)nullptr, d_out_pool, lock);
                NQM_HIP(ctx, hipGetLastError());
            }
            if (per_col_out) {
                std::vector<uint64_t> pcs(32, 0);
                NQM_HIP(ctx, hipMemcpyAsync(pcs.data(), g.per_col.ptr, 32 * 8, hipMemcpyDeviceToHost, st));
                NQM_HIP(ctx, hipStreamSynchronize(st));
                std::memcpy(per_col_out, pcs.data(), n * sizeof(uint64_t));
            }
            NQM_HIP(ctx, hipStreamSynchronize(st));
            *count_out = next;
            return NQM_OK;
        }
        if ((rc = ensure_dev(ctx, g.lvl[cur ^ 1], std::max<uint64_t>(next, 1) * sizeof(GenItem))) != NQM_OK)
            return rc;
        hipLaunchKernelGGL(gen_write_kernel, dim3(nblocks), dim3(kGenThreads), 0, st, in, (uint32_t)count,
                           mask, row, static_cast<uint32_t*>(g.sums.ptr),
                           static_cast<GenItem*>(g.lvl[cur ^ 1].ptr), (uint8_t*)nullptr, 0u);
        NQM_HIP(ctx, hipGetLastError());
        count = next;
        cur ^= 1;
    }
    // a level emptied out before reaching `lock`
    *count_out = 0;
    if (per_col_out) std::memset(per_col_out, 0, n * sizeof(uint64_t));
    NQM_HIP(ctx, hipStreamSynchronize(st));
    return NQM_OK;
}

}  // namespace

}  // namespace nqm

using namespace nqm;

extern "C" {

int nqm_pool_count(nqm_ctx* ctx, uint32_t n, uint32_t lock, uint32_t flags, uint64_t* count,
                   uint64_t* per_col) {
    if (!ctx || !count) return NQM_EINVAL;
    if (flags & ~NQM_MIRROR_WEIGHT) return fail(ctx, NQM_EINVAL, "unknown flags 0x%x", flags);
    NQM_HIP(ctx, hipSetDevice(ctx->device));
    return generate(ctx, n, lock, flags & NQM_MIRROR_WEIGHT, nullptr, 0, count, per_col);
}

int nqm_pool_generate_device(nqm_ctx* ctx, uint32_t n, uint32_t lock, uint32_t flags, uint8_t** d_out,
                             uint64_t* count) {
    if (!ctx || !d_out || !count) return NQM_EINVAL;
    if (flags & ~NQM_MIRROR_WEIGHT) return fail(ctx, NQM_EINVAL, "unknown flags 0x%x", flags);
    NQM_HIP(ctx, hipSetDevice(ctx->device));
    *d_out = nullptr;
    uint64_t cnt = 0;
    int rc = generate(ctx, n, lock, flags & NQM_MIRROR_WEIGHT, nullptr, 0, &cnt, nullptr);
    if (rc != NQM_OK) return rc;
    void* p = nullptr;
    if ((rc = nqm_dev_alloc(ctx, std::max<uint64_t>(cnt * lock, 1), &p)) != NQM_OK) return rc;
    rc = generate(ctx, n, lock, flags & NQM_MIRROR_WEIGHT, static_cast<uint8_t*>(p), cnt, &cnt, nullptr);
    if (rc != NQM_OK) {
        (void)nqm_dev_free(ctx, p);
        return rc;
    }
    *d_out = static_cast<uint8_t*>(p);
    *count = cnt;
    return NQM_OK;
}

int nqm_pool_generate(nqm_ctx* ctx, uint32_t n, uint32_t lock, uint32_t flags, uint8_t* out, uint64_t cap,
                      uint64_t* count) {
    if (!ctx || !count || (!out && cap)) return NQM_EINVAL;
    uint8_t* d = nullptr;
    uint64_t cnt = 0;
    int rc = nqm_pool_generate_device(ctx, n, lock, flags, &d, &cnt);
    if (rc != NQM_OK) return rc;
    *count = cnt;
    if (cnt > cap) {
        (void)nqm_dev_free(ctx, d);
        return fail(ctx, NQM_ERANGE, "pool of %llu prefixes exceeds capacity %llu", (unsigned long long)cnt,
                    (unsigned long long)cap);
    }
    rc = nqm_memcpy_d2h(ctx, out, d, cnt * lock);
    (void)nqm_dev_free(ctx, d);
    return rc;
}

// gen_states_threaded, n_queens.cu:128-219, on device pool counts: start at
// floor(ln k / ln n) (:137-139); keep the previous pool while the next one is
// not smaller (downslope bound, deffinitions.cuh:57, :188-189), lock <= n and
// the pool <= 1.2*k (:135, :191); a row-0 column whose pool reaches the cap is
// the generator's ERANGE (:76-80), which ends the search (:177-180). The
// reference keeps the previous iteration's pool (:196-217).
int nqm_choose_lock(nqm_ctx* ctx, uint32_t n, uint64_t how_many, uint32_t flags, uint32_t* lock_out,
                    uint64_t* count_out) {
    if (!ctx || !lock_out || !count_out) return NQM_EINVAL;
    if (n < 2 || n > kMaxN || how_many == 0) return fail(ctx, NQM_EINVAL, "n=%u how_many=%llu", n,
                                                          (unsigned long long)how_many);
    const uint64_t cap = (uint64_t)((double)how_many * (1.0 + 0.2));
    unsigned lock = (unsigned)(std::log((double)how_many) / std::log((double)n));
    if (lock >= n) lock = n - 1;
    if (lock == 0) return fail(ctx, NQM_ERANGE, "request of %llu states is below one row for n=%u",
                               (unsigned long long)how_many, n);
    bool have_actual = false, have_tmp = false;
    uint64_t actual = 0, tmp = 0;
    unsigned actual_lock = 0, tmp_lock = 0;
    std::vector<uint64_t> per_col(n);
    for (;;) {
        actual = tmp;
        actual_lock = tmp_lock;
        have_actual = have_tmp;
        have_tmp = false;
        tmp = 0;
        if (lock >= n) break;
        uint64_t total = 0;
        int rc = nqm_pool_count(ctx, n, lock, flags, &total, per_col.data());
        if (rc != NQM_OK) return rc;
        bool err = false;
        for (uint32_t c = 0; c < n; ++c)
            if (per_col[c] >= cap) err = true;
        if (err) break;
        tmp = total;
        tmp_lock = lock;
        have_tmp = true;
        ++lock;
        if (!(tmp >= actual && lock <= n && tmp <= cap)) break;
    }
    if (!have_actual) return fail(ctx, NQM_ERANGE, "no lock row gives a pool within 1.2 x %llu",
                                  (unsigned long long)how_many);
    *lock_out = actual_lock;
    *count_out = actual;
    return NQM_OK;
}

}  // extern "C"
