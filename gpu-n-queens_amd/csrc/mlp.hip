// mlp.hip -- subtree-size predictor: 12 -> 64 -> 64 -> 1 MLP, ReLU, FP16
// weights, FP32 accumulation, trained on the device by its own kernels.
//
// No reference counterpart (the reference has no predictor). It estimates
// log2(nodes below a queue item + 1) from relative next-row bitsets, so the
// batch path can split the largest items deeper and queue them first.
//
// Features (16 halves per item, 12 used): depth/16; popcounts of the free
// columns, of the down-left and of the down-right attacked columns at the next
// row (each /16); the free-cell count of rows r..r+7 left by the queens
// already placed (each /16, 0 past row n-1).
//
// MFMA layout (v_mfma_f32_16x16x16f16, wave64): lane l holds
//   A[l%16][4(l/16)+j], B[4(l/16)+j][l%16], D[4(l/16)+j][l%16], j = 0..3.
// The forward pass is computed transposed (hidden x samples): a layer's D
// tile (hidden rows, sample columns) is then exactly the next layer's B
// operand, so activations never leave registers. The weight-gradient GEMMs
// reduce over samples and stage their operands through wave-private LDS.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <vector>

#include "nqm_internal.h"
#include "nqm_items.h"

namespace nqm {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int kIn = 16;   // 12 features + 4 zero pads
constexpr int kHid = 64;
constexpr int kParams = kIn * kHid + kHid + kHid * kHid + kHid + kHid + 1;
// parameter offsets (floats)
constexpr int kW1 = 0;                    // [kIn][kHid]
constexpr int kB1 = kW1 + kIn * kHid;     // [kHid]
constexpr int kW2 = kB1 + kHid;           // [kHid in][kHid out]
constexpr int kB2 = kW2 + kHid * kHid;    // [kHid]
constexpr int kW3 = kB2 + kHid;           // [kHid]
constexpr int kB3 = kW3 + kHid;           // scalar

__device__ __forceinline__ float4_t mfma16(half4_t a, half4_t b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------
__device__ __forceinline__ void item_features(uint32_t n, uint32_t row, uint32_t cols, uint32_t ld,
                                              uint32_t rd, _Float16* out) {
    const uint32_t mask = (1u << n) - 1u;
    const float s = 1.0f / 16.0f;
    out[0] = (_Float16)(row * s);
    out[1] = (_Float16)(__popc(~cols & mask) * s);
    out[2] = (_Float16)(__popc(ld & mask) * s);
    out[3] = (_Float16)(__popc(rd & mask) * s);
    for (uint32_t j = 0; j < 8; ++j) {
        uint32_t c = 0;
        if (row + j < n) {
            const uint32_t att = cols | ((ld << j) & mask) | (rd >> j);
            c = __popc(~att & mask);
        }
        out[4 + j] = (_Float16)(c * s);
    }
    for (int j = 12; j < kIn; ++j) out[j] = (_Float16)0.0f;
}

__global__ void mlp_features_pool_kernel(const uint8_t* __restrict__ pool, uint64_t count, uint32_t n,
                                         uint32_t depth, _Float16* __restrict__ feats) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t mask = (1u << n) - 1u;
    uint32_t cols = 0, ld = 0, rd = 0;
    for (uint32_t r = 0; r < depth; ++r) {
        const uint32_t b = 1u << (pool[i * depth + r] & 31u);
        cols |= b;
        ld = ((ld | b) << 1) & mask;
        rd = (rd | b) >> 1;
    }
    item_features(n, depth, cols, ld, rd, feats + i * kIn);
}

__global__ void mlp_features_items_kernel(const Item* __restrict__ items, const uint32_t* __restrict__ count,
                                          uint32_t max_count, _Float16* __restrict__ feats) {
    const uint32_t c = min(*count, max_count);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < c; i += gridDim.x * blockDim.x) {
        const Item it = items[i];
        item_features(meta_n(it.meta), meta_row(it.meta), it.cols, it.ld, it.rd, feats + (size_t)i * kIn);
    }
}

// ---------------------------------------------------------------------------
// forward (+ backward) over 16-sample tiles, one wave per tile stream
// ---------------------------------------------------------------------------
struct MlpArgs {
    const _Float16* wh;      // fp16 weights (kParams, same offsets)
    const float* wf;         // fp32 biases read from here (same offsets)
    const _Float16* x;       // [count][kIn]
    const float* t;          // labels (backward only)
    const uint32_t* index;   // optional sample permutation
    uint32_t count;          // samples in this call (an upper bound when count_dev is set)
    const uint32_t* count_dev;  // optional: the true sample count, on the device
    float inv_batch;         // 1 / samples contributing to the mean loss
    float* y;                // optional predictions [count]
    float* grad;             // fp32 gradient accumulators (kParams), backward only
    float* loss;             // sum of squared errors, backward only
    int backward;
};

constexpr int kMlpWaves = 4;  // waves per workgroup

// kBackward = false: prediction only; the gradient state is compiled out,
// which keeps the kernel small enough for several waves per SIMD.
template <bool kBackward>
__global__ void __launch_bounds__(64 * kMlpWaves) mlp_pass_kernel(MlpArgs a) {
    // wave-private LDS: h1 [64][16], dh2 [64][16], dh1 [64][16] as f16
    __shared__ _Float16 lds_h1[kMlpWaves][kHid * 16];
    __shared__ _Float16 lds_dh2[kMlpWaves][kHid * 16];
    __shared__ _Float16 lds_dh1[kMlpWaves][kHid * 16];
    __shared__ float lds_x[kMlpWaves][16 * kIn];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, q4 = (lane >> 4) * 4;
    const _Float16* W1 = a.wh + kW1;
    const _Float16* W2 = a.wh + kW2;
    const float* B1 = a.wf + kB1;
    const float* B2 = a.wf + kB2;
    const float* W3 = a.wf + kW3;
    const float b3 = a.wf[kB3];

    // weight operands that do not depend on the tile
    half4_t A1[4], A2[4][4], A2t[4][4];
    for (int mt = 0; mt < 4; ++mt) {
        for (int j = 0; j < 4; ++j) A1[mt][j] = W1[(q4 + j) * kHid + 16 * mt + r16];  // W1^T
        for (int kt = 0; kt < 4; ++kt)
            for (int j = 0; j < 4; ++j) {
                A2[mt][kt][j] = W2[(16 * kt + q4 + j) * kHid + 16 * mt + r16];  // W2^T
                // W2 itself: only the backward pass multiplies by it
                A2t[mt][kt][j] = kBackward ? W2[(16 * mt + r16) * kHid + 16 * kt + q4 + j] : (_Float16)0.0f;
            }
    }
    float4_t gW2[4][4], gW1[4];
    float gb1[16], gb2[16], gw3[16], gb3 = 0.f, lsum = 0.f;
    for (int i = 0; i < 4; ++i) {
        gW1[i] = float4_t{0, 0, 0, 0};
        for (int k = 0; k < 4; ++k) gW2[i][k] = float4_t{0, 0, 0, 0};
    }
    for (int i = 0; i < 16; ++i) gb1[i] = gb2[i] = gw3[i] = 0.f;

    const uint32_t count = a.count_dev ? min(a.count, *a.count_dev) : a.count;
    const uint32_t ntiles = (count + 15u) / 16u;
    const uint32_t gw = blockIdx.x * kMlpWaves + wave, nw = gridDim.x * kMlpWaves;
    for (uint32_t tile = gw; tile < ntiles; tile += nw) {
        const uint32_t s = tile * 16u + (uint32_t)r16;  // this lane's sample (column)
        const bool valid = s < count;
        const uint32_t src = valid ? (a.index ? a.index[s] : s) : 0u;
        half4_t bx;
        for (int j = 0; j < 4; ++j) bx[j] = valid ? a.x[(size_t)src * kIn + q4 + j] : (_Float16)0.0f;

        // layer 1: h1 = relu(W1^T x + b1), D layout (hidden, sample)
        float4_t h1[4];
        half4_t h1h[4];
        for (int mt = 0; mt < 4; ++mt) {
            float4_t c;
            for (int j = 0; j < 4; ++j) c[j] = B1[16 * mt + q4 + j];
            c = mfma16(A1[mt], bx, c);
            for (int j = 0; j < 4; ++j) {
                h1[mt][j] = c[j] > 0.f ? c[j] : 0.f;
                h1h[mt][j] = (_Float16)h1[mt][j];
            }
        }
        // layer 2
        float4_t h2[4];
        for (int mt = 0; mt < 4; ++mt) {
            float4_t c;
            for (int j = 0; j < 4; ++j) c[j] = B2[16 * mt + q4 + j];
            for (int kt = 0; kt < 4; ++kt) c = mfma16(A2[mt][kt], h1h[kt], c);
            for (int j = 0; j < 4; ++j) h2[mt][j] = c[j] > 0.f ? c[j] : 0.f;
        }
        // layer 3: y = w3 . h2 + b3 (sum over the 4 lanes sharing a sample)
        float part = 0.f;
        for (int mt = 0; mt < 4; ++mt)
            for (int j = 0; j < 4; ++j) part += W3[16 * mt + q4 + j] * h2[mt][j];
        part += __shfl_xor(part, 16, 64);
        part += __shfl_xor(part, 32, 64);
        const float y = part + b3;
        if (a.y && valid && lane < 16) a.y[s] = y;
        if (!kBackward) continue;

        // ---- backward ----
        const float err = valid ? (y - a.t[src]) : 0.f;
        if (lane < 16) lsum += err * err;
        const float dy = 2.f * err * a.inv_batch;
        if (lane < 16) gb3 += dy;
        float4_t dh2[4];
        for (int mt = 0; mt < 4; ++mt)
            for (int j = 0; j < 4; ++j) {
                const int h = 16 * mt + q4 + j;
                gw3[4 * mt + j] += dy * h2[mt][j];
                const float g = h2[mt][j] > 0.f ? dy * W3[h] : 0.f;
                dh2[mt][j] = g;
                gb2[4 * mt + j] += g;
            }
        // dh1 = (W2 dh2) * relu'(h1): A = W2 (in x out), B = dh2 (out x sample)
        half4_t dh2h[4];
        for (int kt = 0; kt < 4; ++kt)
            for (int j = 0; j < 4; ++j) dh2h[kt][j] = (_Float16)dh2[kt][j];
        float4_t dh1[4];
        for (int mt = 0; mt < 4; ++mt) {
            float4_t c = float4_t{0, 0, 0, 0};
            for (int kt = 0; kt < 4; ++kt) c = mfma16(A2t[mt][kt], dh2h[kt], c);
            for (int j = 0; j < 4; ++j) {
                dh1[mt][j] = h1[mt][j] > 0.f ? c[j] : 0.f;
                gb1[4 * mt + j] += dh1[mt][j];
            }
        }
        // stage h1, dh2, dh1 (hidden-major, 16 samples per row) and x
        _Float16* sh1 = lds_h1[wave];
        _Float16* sdh2 = lds_dh2[wave];
        _Float16* sdh1 = lds_dh1[wave];
        float* sx = lds_x[wave];
        for (int mt = 0; mt < 4; ++mt)
            for (int j = 0; j < 4; ++j) {
                const int h = 16 * mt + q4 + j;
                sh1[h * 16 + r16] = h1h[mt][j];
                sdh2[h * 16 + r16] = dh2h[mt][j];
                sdh1[h * 16 + r16] = (_Float16)dh1[mt][j];
            }
        for (int j = 0; j < 4; ++j) sx[r16 * kIn + q4 + j] = (float)bx[j];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // dW2[in][out] += h1[in][s] dh2[out][s]: A = h1 (in x s), B = dh2^T (s x out)
        for (int mt = 0; mt < 4; ++mt) {
            half4_t av;
            for (int j = 0; j < 4; ++j) av[j] = sh1[(16 * mt + r16) * 16 + q4 + j];
            for (int nt = 0; nt < 4; ++nt) {
                half4_t bv;
                for (int j = 0; j < 4; ++j) bv[j] = sdh2[(16 * nt + r16) * 16 + q4 + j];
                gW2[mt][nt] = mfma16(av, bv, gW2[mt][nt]);
            }
        }
        // dW1[f][h] += x[s][f] dh1[h][s]: A = x^T (f x s), B = dh1^T (s x h)
        {
            half4_t av;
            for (int j = 0; j < 4; ++j) av[j] = (_Float16)sx[(q4 + j) * kIn + r16];
            for (int nt = 0; nt < 4; ++nt) {
                half4_t bv;
                for (int j = 0; j < 4; ++j) bv[j] = sdh1[(16 * nt + r16) * 16 + q4 + j];
                gW1[nt] = mfma16(av, bv, gW1[nt]);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (!kBackward) return;

    // fold this wave's accumulators into the global gradient
    for (int mt = 0; mt < 4; ++mt)
        for (int nt = 0; nt < 4; ++nt)
            for (int j = 0; j < 4; ++j)
                atomicAdd(&a.grad[kW2 + (16 * mt + q4 + j) * kHid + 16 * nt + r16], gW2[mt][nt][j]);
    for (int nt = 0; nt < 4; ++nt)
        for (int j = 0; j < 4; ++j) atomicAdd(&a.grad[kW1 + (q4 + j) * kHid + 16 * nt + r16], gW1[nt][j]);
    for (int i = 0; i < 16; ++i) {  // per-hidden sums over the 16 sample lanes
        float v1 = gb1[i], v2 = gb2[i], v3 = gw3[i];
        for (int o = 1; o < 16; o <<= 1) {
            v1 += __shfl_xor(v1, o, 64);
            v2 += __shfl_xor(v2, o, 64);
            v3 += __shfl_xor(v3, o, 64);
        }
        if (r16 == 0) {
            const int h = 16 * (i / 4) + q4 + (i % 4);
            atomicAdd(&a.grad[kB1 + h], v1);
            atomicAdd(&a.grad[kB2 + h], v2);
            atomicAdd(&a.grad[kW3 + h], v3);
        }
    }
    for (int o = 1; o < 16; o <<= 1) {
        gb3 += __shfl_xor(gb3, o, 64);
        lsum += __shfl_xor(lsum, o, 64);
    }
    if (lane == 0) {
        atomicAdd(&a.grad[kB3], gb3);
        atomicAdd(a.loss, lsum);
    }
}

// Adam on fp32 master weights; writes the fp16 copy used by the passes.
__global__ void mlp_adam_kernel(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                                float* __restrict__ g, _Float16* __restrict__ wh, float lr, float b1c,
                                float b2c, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / b1c) / (sqrtf(vi / b2c) + 1e-8f);
    w[i] = wi;
    wh[i] = (_Float16)wi;
    g[i] = 0.f;
}

__global__ void mlp_to_half_kernel(const float* __restrict__ w, _Float16* __restrict__ wh, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wh[i] = (_Float16)w[i];
}

}  // namespace nqm

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct nqm_mlp {
    nqm_ctx* ctx = nullptr;
    nqm::DevBuf w, m, v, g, wh, loss;
    int step = 0;
    uint64_t version = 0;  // new value from g_mlp_version at every weight change
};

namespace nqm {

// Weight versions: every change of any model's weights takes the next value,
// so that a batch plan keyed by (model, version) never outlives the weights
// it was planned with, even if a new model reuses a freed model's address.
static std::atomic<uint64_t> g_mlp_version{0};

static void mlp_touch(nqm_mlp* m) { m->version = g_mlp_version.fetch_add(1, std::memory_order_relaxed) + 1; }

uint64_t mlp_version(const nqm_mlp* m) { return m->version; }

static uint64_t mix64(uint64_t& s) {  // splitmix64
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static int mlp_launch_pass(nqm_mlp* mlp, const _Float16* x, const float* t, const uint32_t* index,
                           uint32_t count, float inv_batch, float* y, bool backward,
                           const uint32_t* count_dev = nullptr) {
    nqm_ctx* ctx = mlp->ctx;
    MlpArgs a;
    a.wh = static_cast<const _Float16*>(mlp->wh.ptr);
    a.wf = static_cast<const float*>(mlp->w.ptr);
    a.x = x;
    a.t = t;
    a.index = index;
    a.count = count;
    a.count_dev = count_dev;
    a.inv_batch = inv_batch;
    a.y = y;
    a.grad = static_cast<float*>(mlp->g.ptr);
    a.loss = static_cast<float*>(mlp->loss.ptr);
    a.backward = backward ? 1 : 0;
    const uint32_t tiles = (count + 15u) / 16u;
    // backward: a handful of tiles per wave keeps the atomic fold cheap;
    // forward: enough waves to fill the chip, each loading the weights once
    // for many tiles
    const uint32_t fwd_waves = (uint32_t)ctx->prop.multiProcessorCount * 32u;
    const uint32_t waves = backward ? std::max(1u, std::min(tiles, tiles / 8u + 1u))
                                    : std::max(1u, std::min(tiles, fwd_waves));
    const uint32_t blocks = std::max(1u, std::min(65535u, (waves + kMlpWaves - 1) / kMlpWaves));
    if (backward)
        hipLaunchKernelGGL(mlp_pass_kernel<true>, dim3(blocks), dim3(64 * kMlpWaves), 0, ctx->stream, a);
    else
        hipLaunchKernelGGL(mlp_pass_kernel<false>, dim3(blocks), dim3(64 * kMlpWaves), 0, ctx->stream, a);
    NQM_HIP(ctx, hipGetLastError());
    return NQM_OK;
}

int mlp_predict_device(nqm_mlp* mlp, const void* d_x, uint32_t count, float* d_y, const uint32_t* d_count) {
    if (count == 0) return NQM_OK;
    return mlp_launch_pass(mlp, static_cast<const _Float16*>(d_x), nullptr, nullptr, count, 0.f, d_y, false,
                           d_count);
}

hipError_t mlp_features_items(hipStream_t s, const Item* items, const uint32_t* count, void* feats,
                              uint32_t max_count) {
    const uint32_t blocks = std::max(1u, std::min(4096u, (max_count + 255u) / 256u));
    hipLaunchKernelGGL(mlp_features_items_kernel, dim3(blocks), dim3(256), 0, s, items, count, max_count,
                       static_cast<_Float16*>(feats));
    return hipGetLastError();
}

}  // namespace nqm

using namespace nqm;

extern "C" {

uint32_t nqm_mlp_param_count(void) { return (uint32_t)kParams; }
uint32_t nqm_mlp_feature_width(void) { return (uint32_t)kIn; }

int nqm_mlp_create(nqm_ctx* ctx, uint64_t seed, nqm_mlp** out) {
    if (!ctx || !out) return NQM_EINVAL;
    *out = nullptr;
    NQM_HIP(ctx, hipSetDevice(ctx->device));
    nqm_mlp* m = new (std::nothrow) nqm_mlp();
    if (!m) return NQM_ENOMEM;
    m->ctx = ctx;
    int rc;
    for (DevBuf* d : {&m->w, &m->m, &m->v, &m->g})
        if ((rc = ensure_dev(ctx, *d, kParams * sizeof(float))) != NQM_OK) {
            nqm_mlp_destroy(m);
            return rc;
        }
    if ((rc = ensure_dev(ctx, m->wh, kParams * sizeof(_Float16))) != NQM_OK ||
        (rc = ensure_dev(ctx, m->loss, 16)) != NQM_OK) {
        nqm_mlp_destroy(m);
        return rc;
    }
    // He-uniform init from splitmix64; pad-feature rows of W1 stay 0
    std::vector<float> w(kParams, 0.f);
    uint64_t s = seed;
    auto uni = [&](float lim) { return ((mix64(s) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0) * lim; };
    for (int f = 0; f < 12; ++f)
        for (int h = 0; h < kHid; ++h) w[kW1 + f * kHid + h] = (float)uni(std::sqrt(6.0 / 12.0));
    for (int i = 0; i < kHid * kHid; ++i) w[kW2 + i] = (float)uni(std::sqrt(6.0 / kHid));
    for (int h = 0; h < kHid; ++h) w[kW3 + h] = (float)uni(std::sqrt(3.0 / kHid));
    w[kB3] = 10.0f;  // log2 subtree sizes are O(10)
    rc = nqm_mlp_set_params(m, w.data());
    if (rc != NQM_OK) {
        nqm_mlp_destroy(m);
        return rc;
    }
    *out = m;
    return NQM_OK;
}

void nqm_mlp_destroy(nqm_mlp* m) {
    if (!m) return;
    for (DevBuf* d : {&m->w, &m->m, &m->v, &m->g, &m->wh, &m->loss})
        if (d->ptr) (void)hipFree(d->ptr);
    delete m;
}

int nqm_mlp_set_params(nqm_mlp* m, const float* params) {
    if (!m || !params) return NQM_EINVAL;
    nqm_ctx* ctx = m->ctx;
    hipStream_t st = ctx->stream;
    mlp_touch(m);
    NQM_HIP(ctx, hipMemcpyAsync(m->w.ptr, params, kParams * 4, hipMemcpyHostToDevice, st));
    NQM_HIP(ctx, hipMemsetAsync(m->m.ptr, 0, kParams * 4, st));
    NQM_HIP(ctx, hipMemsetAsync(m->v.ptr, 0, kParams * 4, st));
    NQM_HIP(ctx, hipMemsetAsync(m->g.ptr, 0, kParams * 4, st));
    hipLaunchKernelGGL(mlp_to_half_kernel, dim3((kParams + 255) / 256), dim3(256), 0, st,
                       static_cast<const float*>(m->w.ptr), static_cast<_Float16*>(m->wh.ptr), kParams);
    NQM_HIP(ctx, hipGetLastError());
    NQM_HIP(ctx, hipStreamSynchronize(st));
    m->step = 0;
    return NQM_OK;
}

int nqm_mlp_get_params(nqm_mlp* m, float* params) {
    if (!m || !params) return NQM_EINVAL;
    return nqm_memcpy_d2h(m->ctx, params, m->w.ptr, kParams * 4);
}

int nqm_mlp_features_pool(nqm_ctx* ctx, uint32_t n, uint32_t depth, const uint8_t* d_pool, uint64_t count,
                          void* d_feats) {
    if (!ctx || !d_feats || (count && depth && !d_pool) || n < 1 || n > kMaxN || depth > n) return NQM_EINVAL;
    if (!count) return NQM_OK;
    const uint64_t blocks = (count + 255) / 256;
    hipLaunchKernelGGL(mlp_features_pool_kernel, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream, d_pool,
                       count, n, depth, static_cast<_Float16*>(d_feats));
    NQM_HIP(ctx, hipGetLastError());
    NQM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return NQM_OK;
}

int nqm_mlp_predict(nqm_mlp* m, const void* d_feats, uint32_t count, float* d_out) {
    if (!m || (count && (!d_feats || !d_out))) return NQM_EINVAL;
    int rc = mlp_predict_device(m, d_feats, count, d_out);
    if (rc != NQM_OK) return rc;
    NQM_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
    return NQM_OK;
}

int nqm_mlp_gradient(nqm_mlp* m, const void* d_feats, const float* d_labels, uint32_t count,
                     float* grad_out, float* loss_out) {
    if (!m || !count || !d_feats || !d_labels) return NQM_EINVAL;
    nqm_ctx* ctx = m->ctx;
    NQM_HIP(ctx, hipMemsetAsync(m->g.ptr, 0, kParams * 4, ctx->stream));
    NQM_HIP(ctx, hipMemsetAsync(m->loss.ptr, 0, 4, ctx->stream));
    int rc = mlp_launch_pass(m, static_cast<const _Float16*>(d_feats), d_labels, nullptr, count,
                             1.0f / (float)count, nullptr, true);
    if (rc != NQM_OK) return rc;
    float loss = 0.f;
    if (grad_out)
        NQM_HIP(ctx, hipMemcpyAsync(grad_out, m->g.ptr, kParams * 4, hipMemcpyDeviceToHost, ctx->stream));
    NQM_HIP(ctx, hipMemcpyAsync(&loss, m->loss.ptr, 4, hipMemcpyDeviceToHost, ctx->stream));
    NQM_HIP(ctx, hipMemsetAsync(m->g.ptr, 0, kParams * 4, ctx->stream));
    NQM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (loss_out) *loss_out = loss / (float)count;
    return NQM_OK;
}

int nqm_mlp_train(nqm_mlp* m, const void* d_feats, const float* d_labels, uint32_t count, uint32_t epochs,
                  uint32_t batch, float lr, uint64_t seed, float val_frac, float* train_loss,
                  float* val_loss) {
    if (!m || !count || !d_feats || !d_labels || !batch || val_frac < 0.f || val_frac >= 1.f)
        return NQM_EINVAL;
    nqm_ctx* ctx = m->ctx;
    hipStream_t st = ctx->stream;
    mlp_touch(m);
    // seeded split: a permutation; the first (1 - val_frac) part trains
    std::vector<uint32_t> perm(count);
    for (uint32_t i = 0; i < count; ++i) perm[i] = i;
    uint64_t s = seed;
    for (uint32_t i = count - 1; i > 0; --i) std::swap(perm[i], perm[mix64(s) % (i + 1)]);
    const uint32_t ntrain = std::max(1u, (uint32_t)((double)count * (1.0 - val_frac)));
    const uint32_t nval = count - ntrain;
    DevBuf d_idx;
    int rc = ensure_dev(ctx, d_idx, (size_t)count * 4);
    if (rc != NQM_OK) return rc;
    struct Free {
        DevBuf& b;
        ~Free() {
            if (b.ptr) (void)hipFree(b.ptr);
        }
    } fr{d_idx};
    uint32_t* idx = static_cast<uint32_t*>(d_idx.ptr);
    std::vector<uint32_t> order(ntrain);
    float last_train = 0.f;
    for (uint32_t e = 0; e < epochs; ++e) {
        for (uint32_t i = 0; i < ntrain; ++i) order[i] = perm[i];
        for (uint32_t i = ntrain - 1; i > 0; --i) std::swap(order[i], order[mix64(s) % (i + 1)]);
        NQM_HIP(ctx, hipMemcpyAsync(idx, order.data(), (size_t)ntrain * 4, hipMemcpyHostToDevice, st));
        NQM_HIP(ctx, hipMemsetAsync(m->loss.ptr, 0, 4, st));
        for (uint32_t b0 = 0; b0 < ntrain; b0 += batch) {
            const uint32_t nb = std::min(batch, ntrain - b0);
            rc = mlp_launch_pass(m, static_cast<const _Float16*>(d_feats), d_labels, idx + b0, nb,
                                 1.0f / (float)nb, nullptr, true);
            if (rc != NQM_OK) return rc;
            ++m->step;
            const float b1c = 1.f - std::pow(0.9f, (float)m->step);
            const float b2c = 1.f - std::pow(0.999f, (float)m->step);
            hipLaunchKernelGGL(mlp_adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, st,
                               static_cast<float*>(m->w.ptr), static_cast<float*>(m->m.ptr),
                               static_cast<float*>(m->v.ptr), static_cast<float*>(m->g.ptr),
                               static_cast<_Float16*>(m->wh.ptr), lr, b1c, b2c, kParams);
            NQM_HIP(ctx, hipGetLastError());
        }
        float l = 0.f;
        NQM_HIP(ctx, hipMemcpyAsync(&l, m->loss.ptr, 4, hipMemcpyDeviceToHost, st));
        NQM_HIP(ctx, hipStreamSynchronize(st));
        last_train = l / (float)ntrain;
    }
    if (train_loss) *train_loss = last_train;
    if (val_loss) {
        *val_loss = 0.f;
        if (nval) {
            // validation MSE with a forward+loss pass that does not touch weights
            NQM_HIP(ctx, hipMemcpyAsync(idx, perm.data() + ntrain, (size_t)nval * 4, hipMemcpyHostToDevice, st));
            NQM_HIP(ctx, hipMemsetAsync(m->loss.ptr, 0, 4, st));
            rc = mlp_launch_pass(m, static_cast<const _Float16*>(d_feats), d_labels, idx, nval, 0.f, nullptr,
                                 true);
            if (rc != NQM_OK) return rc;
            float l = 0.f;
            NQM_HIP(ctx, hipMemcpyAsync(&l, m->loss.ptr, 4, hipMemcpyDeviceToHost, st));
            NQM_HIP(ctx, hipMemsetAsync(m->g.ptr, 0, kParams * 4, st));
            NQM_HIP(ctx, hipStreamSynchronize(st));
            *val_loss = l / (float)nval;
        }
    }
    return NQM_OK;
}

}  // extern "C"
