// nqm_static.hip -- the counting entry point that allocates nothing.
//
// nqm_count_pool_device_static runs the static-pool count (K1) on a pool
// already in device memory, with every byte of scratch it needs (the
// per-workgroup partials and the invalid-prefix flag) in `scratch`, a
// file-scope device array of fixed size: no hipMalloc, no hipFree, no
// memset, no host synchronization. It is for callers that cannot allocate
// (code under HIP graph capture, tight serving loops). Results land in the
// caller's device buffer in stream order.
//
// Each GPU the library runs on holds its own copy of `scratch` until the
// process ends. Host code never passes the array's name to a kernel (in host
// code it is a host-side stand-in, not a device address): the address of
// each device's copy comes from hipGetSymbolAddress, called once per device
// with that device current, when the first context on the device is
// created (nqm_create -> static_scratch_init), and kept per device.
//
// Every call on one GPU uses that GPU's copy. Calls that overlap on one GPU
// (two streams, two host threads) race on it; the caller serializes them.
// Each call clears what must start at zero (the flag) with a fill kernel
// before the count, and reads only the records its own workgroups wrote.
#include <mutex>

#include "nqm_internal.h"

namespace nqm {

namespace {

// Largest K1 launch the scratch holds: 2^20 one-wave workgroups, i.e. pools
// of up to 2^26 = 67,108,864 prefixes (the reference's default N=18 pool has
// 37,956,796).
constexpr uint32_t kStaticMaxRecords = 1u << 20;
// u64 words: a 256-byte header (word 0: the invalid-prefix flag), then two
// words {solutions, nodes} per record.
constexpr size_t kStaticHeaderWords = 32;
constexpr size_t kStaticWords = kStaticHeaderWords + 2 * (size_t)kStaticMaxRecords;  // 16 MiB + 256 B

__device__ __align__(256) uint64_t scratch[kStaticWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint64_t* g_scratch[kMaxDevices] = {};

__global__ void static_clear_kernel(uint64_t* __restrict__ header) {
    if (threadIdx.x < kStaticHeaderWords) header[threadIdx.x] = 0;
}

// One block: the records of a K1 launch of nblocks workgroups into
// result[0] = solutions, result[1] = nodes, result[2] = the flag (nonzero:
// an invalid prefix). Fixed summation order: the same inputs give the same
// sums on every replay.
__global__ void __launch_bounds__(256)
static_sum_kernel(const uint64_t* __restrict__ buf, uint32_t nblocks, uint64_t* __restrict__ result) {
    __shared__ uint64_t part[2][4];
    const uint64_t* const recs = buf + kStaticHeaderWords;
    uint64_t s = 0, nd = 0;
    for (uint32_t b = threadIdx.x; b < nblocks; b += 256u) {
        const uint64_t* r = recs + 2u * k1_part_slot(b, nblocks);
        s += r[0];
        nd += r[1];
    }
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        nd += __shfl_xor(nd, o, 64);
    }
    if ((threadIdx.x & 63u) == 0) {
        part[0][threadIdx.x >> 6] = s;
        part[1][threadIdx.x >> 6] = nd;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        result[0] = part[0][0] + part[0][1] + part[0][2] + part[0][3];
        result[1] = part[1][0] + part[1][1] + part[1][2] + part[1][3];
        result[2] = buf[0];
    }
}

}  // namespace

// Looks up this device's copy of `scratch` (device `device` current).
int static_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return NQM_EUNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return NQM_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return NQM_EDEVICE;
    }
    g_scratch[device] = static_cast<uint64_t*>(p);
    return NQM_OK;
}

}  // namespace nqm

using namespace nqm;

extern "C" int nqm_count_pool_device_static(nqm_ctx* ctx, uint32_t n, uint32_t depth, const uint8_t* d_pool,
                                            uint64_t count, uint32_t flags, uint64_t* d_result) {
    if (!ctx) return NQM_EINVAL;
    if (n < 1 || n > kMaxN) return fail(ctx, NQM_EINVAL, "n=%u outside 1..%u", n, kMaxN);
    if (depth > n) return fail(ctx, NQM_EINVAL, "depth=%u > n=%u", depth, n);
    if (count == 0) return fail(ctx, NQM_EINVAL, "empty pool");
    if (flags & ~NQM_MIRROR_WEIGHT) return fail(ctx, NQM_EINVAL, "unknown flags 0x%x", flags);
    if ((flags & NQM_MIRROR_WEIGHT) && depth == 0)
        return fail(ctx, NQM_EINVAL, "mirror weighting needs a row-0 queen (depth >= 1)");
    if (!d_result || (!d_pool && depth > 0)) return fail(ctx, NQM_EINVAL, "NULL argument");
    K1Launch plan;
    int rc = k1_plan(ctx, n, depth, count, &plan);
    if (rc != NQM_OK) return fail(ctx, rc, "no launch shape for n=%u depth=%u", n, depth);
    if (k1_part_records(plan.blocks) > kStaticMaxRecords)
        return fail(ctx, NQM_ERANGE,
                    "pool of %llu prefixes needs %u workgroups; the allocation-free count holds %u "
                    "(%llu prefixes)",
                    (unsigned long long)count, plan.blocks, kStaticMaxRecords,
                    (unsigned long long)kStaticMaxRecords * plan.threads);
    if (ctx->device < 0 || ctx->device >= kMaxDevices || !g_scratch[ctx->device])
        return fail(ctx, NQM_EDEVICE, "no scratch address for device %d", ctx->device);
    uint64_t* const buf = g_scratch[ctx->device];
    NQM_HIP(ctx, hipSetDevice(ctx->device));
    hipLaunchKernelGGL(static_clear_kernel, dim3(1), dim3(64), 0, ctx->stream, buf);
    NQM_HIP(ctx, hipGetLastError());
    hipError_t e = k1_launch(plan, ctx->stream, d_pool, count, n, depth, flags, buf + kStaticHeaderWords, nullptr,
                             nullptr, reinterpret_cast<uint32_t*>(buf));
    if (e != hipSuccess) return fail(ctx, NQM_EDEVICE, "K1 launch failed: %s", hipGetErrorString(e));
    hipLaunchKernelGGL(static_sum_kernel, dim3(1), dim3(256), 0, ctx->stream, buf, plan.blocks, d_result);
    NQM_HIP(ctx, hipGetLastError());
    return NQM_OK;
}
