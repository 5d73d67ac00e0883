// nqm_api.hip -- context management and the counting entry points of nqm.h.
#include <chrono>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "nqm_internal.h"

namespace nqm {

int fail(nqm_ctx* ctx, int status, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->last_error = buf;
    }
    return status;
}

int ensure_dev(nqm_ctx* ctx, DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (b.cap >= bytes) return NQM_OK;
    if (b.ptr) {
        NQM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        NQM_HIP(ctx, hipFree(b.ptr));
        b.ptr = nullptr;
        b.cap = 0;
    }
    size_t want = bytes + bytes / 4;  // grow with headroom
    if (hipMalloc(&b.ptr, want) != hipSuccess) {
        (void)hipGetLastError();
        b.ptr = nullptr;
        return fail(ctx, NQM_ENOMEM, "hipMalloc of %zu bytes failed", want);
    }
    b.cap = want;
    return NQM_OK;
}

int ensure_host(nqm_ctx* ctx, HostBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (b.cap >= bytes) return NQM_OK;
    if (b.ptr) {
        NQM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        NQM_HIP(ctx, hipHostFree(b.ptr));
        b.ptr = nullptr;
        b.cap = 0;
    }
    size_t want = bytes + bytes / 4;
    if (hipHostMalloc(&b.ptr, want, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        b.ptr = nullptr;
        return fail(ctx, NQM_ENOMEM, "hipHostMalloc of %zu bytes failed", want);
    }
    b.cap = want;
    return NQM_OK;
}

static void free_dev(DevBuf& b) {
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.cap = 0;
}

static void free_host(HostBuf& b) {
    if (b.ptr) (void)hipHostFree(b.ptr);
    b.ptr = nullptr;
    b.cap = 0;
}

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Shared body of the static-pool count: pool already on the device.
static int count_device_impl(nqm_ctx* ctx, uint32_t n, uint32_t depth, const uint8_t* d_pool,
                             uint64_t count, uint32_t flags, uint64_t* d_per_sols,
                             uint64_t* d_per_nodes, nqm_stats* st) {
    K1Launch plan;
    int rc = k1_plan(ctx, n, depth, count, &plan);
    if (rc != NQM_OK) return fail(ctx, rc, "no launch shape for n=%u depth=%u", n, depth);
    const size_t recs = k1_part_records(plan.blocks);
    if ((rc = ensure_dev(ctx, ctx->part, recs * 16)) != NQM_OK) return rc;
    if ((rc = ensure_dev(ctx, ctx->err, 8)) != NQM_OK) return rc;
    if ((rc = ensure_host(ctx, ctx->h_part, recs * 16 + 8)) != NQM_OK) return rc;

    NQM_HIP(ctx, hipMemsetAsync(ctx->err.ptr, 0, 4, ctx->stream));
    NQM_HIP(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
    hipError_t e = k1_launch(plan, ctx->stream, d_pool, count, n, depth, flags,
                             static_cast<uint64_t*>(ctx->part.ptr), d_per_sols, d_per_nodes,
                             static_cast<uint32_t*>(ctx->err.ptr));
    if (e != hipSuccess)
        return fail(ctx, NQM_EDEVICE, "K1 launch failed: %s", hipGetErrorString(e));
    NQM_HIP(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));

    uint64_t* h = static_cast<uint64_t*>(ctx->h_part.ptr);
    NQM_HIP(ctx, hipMemcpyAsync(h, ctx->part.ptr, recs * 16, hipMemcpyDeviceToHost, ctx->stream));
    NQM_HIP(ctx, hipMemcpyAsync(h + 2 * recs, ctx->err.ptr, 4, hipMemcpyDeviceToHost, ctx->stream));
    NQM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    NQM_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));

    uint32_t err_flag = 0;
    std::memcpy(&err_flag, h + 2 * recs, 4);
    if (err_flag) return fail(ctx, NQM_EPREFIX, "pool holds an invalid prefix (n=%u depth=%u)", n, depth);

    // Host combine of per-block partials (GPU_NQueens.cu:162-172).
    uint64_t total = 0, nodes = 0;
    for (uint32_t b = 0; b < plan.blocks; ++b) {
        const uint64_t* r = h + 2 * (size_t)k1_part_slot(b, plan.blocks);
        total += r[0];
        nodes += r[1];
    }
    if (st) {
        st->total = total;
        st->nodes = nodes;
        st->kernel_ms = ms;
        st->launches = 1;
        st->blocks = plan.blocks;
        st->search_ms = ms;
    }
    return NQM_OK;
}

// A plain streaming copy: the reference point for the FETCH_SIZE / WRITE_SIZE
// counters (each byte is read once and written once).
__global__ void dev_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

static int check_count_args(nqm_ctx* ctx, uint32_t n, uint32_t depth, uint64_t count,
                            uint32_t flags) {
    if (!ctx) return NQM_EINVAL;
    if (n < 1 || n > kMaxN) return fail(ctx, NQM_EINVAL, "n=%u outside 1..%u", n, kMaxN);
    if (depth > n) return fail(ctx, NQM_EINVAL, "depth=%u > n=%u", depth, n);
    if (count == 0) return fail(ctx, NQM_EINVAL, "empty pool");
    if (flags & ~NQM_MIRROR_WEIGHT) return fail(ctx, NQM_EINVAL, "unknown flags 0x%x", flags);
    if ((flags & NQM_MIRROR_WEIGHT) && depth == 0)
        return fail(ctx, NQM_EINVAL, "mirror weighting needs a row-0 queen (depth >= 1)");
    return NQM_OK;
}

}  // namespace nqm

using namespace nqm;

extern "C" {

const char* nqm_strerror(int status) {
    switch (status) {
        case NQM_OK: return "ok";
        case NQM_EINVAL: return "invalid argument";
        case NQM_EPREFIX: return "invalid prefix";
        case NQM_ENOMEM: return "out of memory";
        case NQM_EDEVICE: return "device error";
        case NQM_ERANGE: return "out of range";
        case NQM_EIO: return "i/o error";
        case NQM_EFORMAT: return "malformed pool file";
        case NQM_EFLAGS: return "pool file flags differ";
        case NQM_EUNSUPPORTED: return "unsupported";
        default: return "unknown status";
    }
}

const char* nqm_last_error(const nqm_ctx* ctx) { return ctx ? ctx->last_error.c_str() : ""; }

int nqm_create(int device, nqm_ctx** out) {
    if (!out) return NQM_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return NQM_EDEVICE;
    }
    if (device < 0 || device >= ndev) return NQM_EINVAL;
    nqm_ctx* ctx = new (std::nothrow) nqm_ctx();
    if (!ctx) return NQM_ENOMEM;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&ctx->prop, device) != hipSuccess) {
        delete ctx;
        return NQM_EDEVICE;
    }
    if (ctx->prop.warpSize != 64) {  // the kernels are written for wave64
        delete ctx;
        return NQM_EDEVICE;
    }
    if (nqm::static_scratch_init(device) != NQM_OK) {  // the allocation-free count's scratch
        delete ctx;
        return NQM_EDEVICE;
    }
    ctx->lds_per_block = (uint32_t)ctx->prop.sharedMemPerBlock;
    ctx->lds_per_cu = (uint32_t)ctx->prop.maxSharedMemoryPerMultiProcessor;
    if (ctx->lds_per_block > 65536) ctx->lds_per_block = 65536;  // static launch limit
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev_start) != hipSuccess || hipEventCreate(&ctx->ev_stop) != hipSuccess ||
        hipEventCreate(&ctx->ev_x0) != hipSuccess || hipEventCreate(&ctx->ev_x1) != hipSuccess) {
        nqm_destroy(ctx);
        return NQM_EDEVICE;
    }
    *out = ctx;
    return NQM_OK;
}

void nqm_destroy(nqm_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_dev(ctx->pool);
    free_dev(ctx->part);
    free_dev(ctx->err);
    nqm::BatchWork& w = ctx->batch;
    nqm::drop_plan(w.plan);
    for (nqm::DevBuf* d : {&w.init, &w.q[0], &w.q[1], &w.counts, &w.head, &w.spill_count, &w.target, &w.res,
                           &w.part, &w.wave_part, &w.cont[0], &w.cont[1], &w.cont_counts, &w.feats, &w.pred,
                           &w.hist, &w.stat, &w.goal})
        free_dev(*d);
    nqm::SymWork& sw = ctx->sym;
    for (nqm::DevBuf* d : {&sw.ctr, &sw.items, &sw.fin, &sw.pend[0], &sw.pend[1], &sw.xpart, &sw.recs, &sw.rec_counts,
                           &sw.dense, &sw.boards, &sw.mult, &sw.wave, &sw.cpart})
        free_dev(*d);
    nqm::FsWork& fw = ctx->fs;
    for (nqm::DevBuf* d : {&fw.ctr, &fw.lvl[0], &fw.lvl[1], &fw.wave}) free_dev(*d);
    free_host(ctx->h_part);
    free_host(sw.h_items);
    if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
    if (ctx->ev_x0) (void)hipEventDestroy(ctx->ev_x0);
    if (ctx->ev_x1) (void)hipEventDestroy(ctx->ev_x1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int nqm_device_info(const nqm_ctx* ctx, uint32_t* cus, uint32_t* wave_size, uint32_t* clock_khz,
                    uint32_t* lds_per_block, uint32_t* lds_per_cu) {
    if (!ctx) return NQM_EINVAL;
    if (cus) *cus = (uint32_t)ctx->prop.multiProcessorCount;
    if (wave_size) *wave_size = (uint32_t)ctx->prop.warpSize;
    if (clock_khz) *clock_khz = (uint32_t)ctx->prop.clockRate;
    if (lds_per_block) *lds_per_block = ctx->lds_per_block;
    if (lds_per_cu) *lds_per_cu = ctx->lds_per_cu;
    return NQM_OK;
}

int nqm_count_pool_device(nqm_ctx* ctx, uint32_t n, uint32_t depth, const uint8_t* d_pool,
                          uint64_t count, uint32_t flags, uint64_t* d_per_sols,
                          uint64_t* d_per_nodes, nqm_stats* stats) {
    int rc = check_count_args(ctx, n, depth, count, flags);
    if (rc != NQM_OK) return rc;
    if (!d_pool && depth > 0) return fail(ctx, NQM_EINVAL, "NULL pool");
    NQM_HIP(ctx, hipSetDevice(ctx->device));
    const double t0 = now_ms();
    nqm_stats st{};
    rc = count_device_impl(ctx, n, depth, d_pool, count, flags, d_per_sols, d_per_nodes, &st);
    st.wall_ms = now_ms() - t0;
    if (rc == NQM_OK && stats) *stats = st;
    return rc;
}

int nqm_count_pool(nqm_ctx* ctx, uint32_t n, uint32_t depth, const uint8_t* pool, uint64_t count,
                   uint32_t flags, uint64_t* total_out) {
    int rc = check_count_args(ctx, n, depth, count, flags);
    if (rc != NQM_OK) return rc;
    if ((!pool && depth > 0) || !total_out) return fail(ctx, NQM_EINVAL, "NULL argument");
    NQM_HIP(ctx, hipSetDevice(ctx->device));
    const size_t bytes = (size_t)count * depth;
    if ((rc = ensure_dev(ctx, ctx->pool, bytes)) != NQM_OK) return rc;
    if (bytes) NQM_HIP(ctx, hipMemcpyAsync(ctx->pool.ptr, pool, bytes, hipMemcpyHostToDevice, ctx->stream));
    nqm_stats st{};
    rc = count_device_impl(ctx, n, depth, static_cast<const uint8_t*>(ctx->pool.ptr), count, flags,
                           nullptr, nullptr, &st);
    if (rc == NQM_OK) *total_out = st.total;
    return rc;
}

int nqm_solver_driver(uint32_t n, const void* states, uint32_t state_cnt, unsigned locked_row_end,
                      const nqm_gpu_config* configs, unsigned config_cnt, uint64_t* total_out) {
    if (!states || !state_cnt || !configs || !config_cnt || !total_out) return NQM_EINVAL;
    if (n < 1 || n > kMaxN || locked_row_end < 1 || locked_row_end >= n) return NQM_EINVAL;
    if (config_cnt != 1) return NQM_EUNSUPPORTED;
    const size_t sz = nqm_state_size(n);
    std::vector<uint8_t> pool((size_t)state_cnt * locked_row_end);
    const uint8_t* s = static_cast<const uint8_t*>(states);
    for (uint32_t i = 0; i < state_cnt; ++i) {
        const uint8_t* st = s + (size_t)i * sz;
        if ((signed char)st[20] != (signed char)locked_row_end) return NQM_EPREFIX;
        std::memcpy(&pool[(size_t)i * locked_row_end], st + 21, locked_row_end);
    }
    nqm_ctx* ctx = nullptr;
    int rc = nqm_create((int)configs[0].device_id, &ctx);
    if (rc != NQM_OK) return rc;
    rc = nqm_count_pool(ctx, n, locked_row_end, pool.data(), state_cnt, NQM_MIRROR_WEIGHT, total_out);
    nqm_destroy(ctx);
    return rc;
}

int nqm_dev_alloc(nqm_ctx* ctx, size_t bytes, void** d_ptr) {
    if (!ctx || !d_ptr) return NQM_EINVAL;
    NQM_HIP(ctx, hipSetDevice(ctx->device));
    if (hipMalloc(d_ptr, bytes ? bytes : 8) != hipSuccess) {
        (void)hipGetLastError();
        *d_ptr = nullptr;
        return fail(ctx, NQM_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
    }
    return NQM_OK;
}

int nqm_dev_free(nqm_ctx* ctx, void* d_ptr) {
    if (!ctx) return NQM_EINVAL;
    if (!d_ptr) return NQM_OK;
    NQM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    NQM_HIP(ctx, hipFree(d_ptr));
    return NQM_OK;
}

int nqm_dev_copy(nqm_ctx* ctx, void* d_dst, const void* d_src, size_t bytes, double* kernel_ms) {
    if (!ctx || (bytes && (!d_dst || !d_src)) || (bytes % 16)) return NQM_EINVAL;
    NQM_HIP(ctx, hipSetDevice(ctx->device));
    NQM_HIP(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
    hipLaunchKernelGGL(dev_copy_kernel, dim3(4096), dim3(256), 0, ctx->stream, static_cast<const uint4*>(d_src),
                       static_cast<uint4*>(d_dst), bytes / 16);
    NQM_HIP(ctx, hipGetLastError());
    NQM_HIP(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
    NQM_HIP(ctx, hipEventSynchronize(ctx->ev_stop));
    float ms = 0.f;
    NQM_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    if (kernel_ms) *kernel_ms = ms;
    return NQM_OK;
}

int nqm_memcpy_h2d(nqm_ctx* ctx, void* d_dst, const void* h_src, size_t bytes) {
    if (!ctx || (bytes && (!d_dst || !h_src))) return NQM_EINVAL;
    if (!bytes) return NQM_OK;
    NQM_HIP(ctx, hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    NQM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return NQM_OK;
}

int nqm_memcpy_d2h(nqm_ctx* ctx, void* h_dst, const void* d_src, size_t bytes) {
    if (!ctx || (bytes && (!h_dst || !d_src))) return NQM_EINVAL;
    if (!bytes) return NQM_OK;
    NQM_HIP(ctx, hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    NQM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return NQM_OK;
}

}  // extern "C"
