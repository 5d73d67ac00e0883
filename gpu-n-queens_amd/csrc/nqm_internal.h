// nqm_internal.h -- host-side internals shared by the nqm translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/nqm.h"

namespace nqm {

// A growable device buffer owned by a context.
struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
};

// Pinned host staging buffer owned by a context.
struct HostBuf {
    void* ptr = nullptr;
    size_t cap = 0;
};

}  // namespace nqm

struct nqm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    hipDeviceProp_t prop{};
    uint32_t lds_per_block = 0;
    uint32_t lds_per_cu = 0;
    nqm::DevBuf pool, part_sols, part_nodes, err;
    nqm::HostBuf h_part;
    std::string last_error;
};

namespace nqm {

// Record a failure on the context and return the status.
int fail(nqm_ctx* ctx, int status, const char* fmt, ...);

// hipError_t -> NQM_EDEVICE with the call text recorded.
#define NQM_HIP(ctx, expr)                                                            \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return ::nqm::fail((ctx), NQM_EDEVICE, "%s failed: %s (%s:%d)", #expr,     \
                               hipGetErrorString(e_), __FILE__, __LINE__);            \
    } while (0)

int ensure_dev(nqm_ctx* ctx, DevBuf& b, size_t bytes);
int ensure_host(nqm_ctx* ctx, HostBuf& b, size_t bytes);

// Largest n this build counts (u32 bitsets), as deffinitions.cuh:116-118.
constexpr uint32_t kMaxN = 31;

// Static-pool counting kernel launcher (K1). Partials: one u64 per block.
struct K1Launch {
    uint32_t blocks = 0;
    uint32_t threads = 0;
    size_t lds_bytes = 0;
};
int k1_plan(const nqm_ctx* ctx, uint32_t n, uint32_t depth, uint64_t count, K1Launch* plan);
hipError_t k1_launch(const K1Launch& plan, hipStream_t stream, const uint8_t* d_pool,
                     uint64_t count, uint32_t n, uint32_t depth, uint32_t flags,
                     uint64_t* d_block_sols, uint64_t* d_block_nodes, uint64_t* d_per_sols,
                     uint64_t* d_per_nodes, uint32_t* d_err);

}  // namespace nqm
