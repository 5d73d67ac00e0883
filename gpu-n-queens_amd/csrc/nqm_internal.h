// nqm_internal.h -- host-side internals shared by the nqm translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/nqm.h"
#include "nqm_limits.h"

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

namespace nqm {
// Device buffers of the batch solver, kept on the context so that repeated
// calls reuse them (they only grow; nqm_destroy frees them).
// The last graph-mode batch, kept instantiated: a call that repeats its
// solves and tuning launches it again. Any other batch call on the context
// drops it (the graph reads the context's buffers and their contents).
struct GraphPlan {
    bool valid = false;
    std::vector<uint8_t> key;  // the solves and the tuning it was built for
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> host_sols, host_nodes;  // placed on the host when it was built
    uint64_t host_exec = 0;
    uint32_t blocks = 0;
    uint64_t lanes = 0;
};
struct BatchWork {
    // res: u64 executed nodes, u32 error bits, u32 pad, then per solve u64
    // solutions and u64 nodes (one copy brings every result back)
    DevBuf init, q[2], counts, head, spill_count, target, res, part, wave_part;
    DevBuf cont[2], cont_counts;     // parked pieces of budget rounds
    DevBuf feats, pred, hist, stat, goal;  // learned splitting (goal: rows-goal target rows)
    GraphPlan plan;
};
void drop_plan(GraphPlan& plan);

// Device buffers of the symmetric drain (k4_sym.hip), kept across calls
// (they only grow), and the expansion sizes of the last n it counted (the
// tree depends on n only, so a repeated n launches the levels without
// counting them first or waiting for the host).
struct SymWork {
    DevBuf ctr, items, fin, pend[2], xpart, recs, rec_counts, dense, boards, mult, wave, cpart;
    uint32_t sizes_n = 0;                         // 0: no sizes kept
    uint32_t host_fin = 0;                        // items the host puts in the final queue
    std::vector<uint32_t> level_in, level_fin;    // per level: pending items in, final items out
    uint64_t chunks = 0;                          // record chunks the last count of sizes_n used
    uint32_t nrec = 0;                            // and its tie records (the same every count of n)
    bool nrec_known = false;
    HostBuf h_items;                              // pinned staging of the host's items
};

// Device buffers of the full-sweep count (k5_fullsweep.hip), kept across
// calls (they only grow): counters, the two level buffers, per-workgroup sums.
struct FsWork {
    DevBuf ctr, lvl[2], wave;
};
}  // namespace nqm

struct nqm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    hipEvent_t ev_x0 = nullptr, ev_x1 = nullptr;  // the batch expansion
    hipDeviceProp_t prop{};
    uint32_t lds_per_block = 0;
    uint32_t lds_per_cu = 0;
    nqm::DevBuf pool, part, err;
    nqm::BatchWork batch;
    nqm::SymWork sym;
    nqm::FsWork fs;
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
double now_ms();


// ---- K4: symmetry-reduced count on a persistent drain (k4_sym.hip) --------
bool k4_supports(uint32_t n);
int k4_count(nqm_ctx* ctx, uint32_t n, uint64_t* unique, uint64_t* total, nqm_stats* stats);

// Looks up the device address of the allocation-free count's scratch array
// for `device` (current), once per device (nqm_static.hip).
int static_scratch_init(int device);

// ---- K1: static pool (k1_static.hip) --------------------------------------
struct K1Launch {
    uint32_t blocks = 0;
    uint32_t threads = 0;
    size_t lds_bytes = 0;
};
int k1_plan(const nqm_ctx* ctx, uint32_t n, uint32_t depth, uint64_t count, K1Launch* plan);
// K1 writes one record {solutions, nodes} (two u64) per workgroup, grouped
// per XCD: workgroup b (of nblocks) goes to record k1_part_slot(b, nblocks)
// of k1_part_records(nblocks).
__host__ __device__ inline uint32_t k1_part_slot(uint32_t b, uint32_t nblocks) {
    return (b & 7u) * ((nblocks + 7u) / 8u) + (b >> 3);
}
__host__ __device__ inline uint32_t k1_part_records(uint32_t nblocks) { return 8u * ((nblocks + 7u) / 8u); }
hipError_t k1_launch(const K1Launch& plan, hipStream_t stream, const uint8_t* d_pool,
                     uint64_t count, uint32_t n, uint32_t depth, uint32_t flags,
                     uint64_t* d_block_part, uint64_t* d_per_sols, uint64_t* d_per_nodes, uint32_t* d_err);

// ---- K2: persistent queue (k2_queue.hip) ----------------------------------
struct Item;
// A search piece parked at the end of a budgeted drain round: its current
// row and its open rows (bottom first), each as f, avail, ld, rd. A later round resumes it exactly where it stopped, so its nodes
// are never counted twice and a round never multiplies the work items.
constexpr uint32_t kContRows = 32;  // more than any stack needs (n <= 31)
struct Cont {
    uint32_t solve, meta, k, pad;  // meta = item_meta(n, 0, ws); k open rows
    uint32_t st[4];                // f, avail, ld, rd of the current row
    uint32_t rows[kContRows][4];
};

struct K2DrainParams {
    const Item* items = nullptr;
    const uint32_t* item_count = nullptr;
    uint32_t* head = nullptr;
    Item* spill = nullptr;
    uint32_t* spill_count = nullptr;
    uint32_t spill_cap = 0;
    uint64_t* solve_sols = nullptr;
    uint64_t* solve_nodes = nullptr;
    uint64_t* exec_nodes = nullptr;
    uint32_t* err = nullptr;
    uint32_t stack_rows = 1;
    uint32_t refill_idle = 16;
    uint32_t next_refill = 0;  // 0: 3 x refill_idle, at most 56
    uint32_t donate = 1;
    uint64_t budget = 0;
    // parked pieces: resumed from cont_in[0, *cont_in_count) (taken through
    // *cont_head), parked into cont_out (count in *cont_out_count) when the
    // budget ends the round
    const Cont* cont_in = nullptr;
    const uint32_t* cont_in_count = nullptr;
    uint32_t* cont_head = nullptr;
    Cont* cont_out = nullptr;
    uint32_t* cont_out_count = nullptr;
    uint32_t cont_cap = 0;
    // 1: refill from the queue before resuming parked pieces (serving: new
    // requests first); 0: parked pieces first
    uint32_t items_first = 0;
    // Queue segments (1 or kHeadSegs), each with its own head at
    // head[s * kHeadStride]; items_first, the restart rounds and the
    // continuations need one segment (the head tells where the unstarted
    // items begin).
    uint32_t nseg = 1;
    // Per-wave totals, 32 bytes per wave of the launch (nullable): each wave
    // stores its counts there and one block adds them to the per-solve
    // counters after the drain, instead of every wave adding them at its end.
    uint64_t* wave_part = nullptr;
};
constexpr uint32_t kHeadSegs = 8;     // one per XCD
constexpr uint32_t kHeadStride = 64;  // u32: one head per 256 bytes
hipError_t k2_fill_u32(hipStream_t s, uint32_t* p, uint32_t n, uint32_t v);
hipError_t k2_fill_u64(hipStream_t s, uint64_t* p, uint32_t n);
// One level of children in two steps, with no same-address atomics (each
// one would serialize the blocks of the level):
//   k2_expand_count: per-block child counts and placements into `part`
//     (kExpandPartBytes), and the level size into *out_count;
//   k2_expand_write: the children, then the placements of the level added
//     to solve_nodes / exec_nodes. Children keep input order, or, with
//     interleave, the blocks' runs of children are interleaved (bit-reversed
//     block order) so that a queue mixes the regions of its input.
// A write needs the count of the same input, grid and rule just before it
// on the stream; a level that is counted but not written adds nothing.
constexpr uint32_t kExpandMaxGrid = 4096;
constexpr size_t kExpandPartBytes = (size_t)kExpandMaxGrid * 32u;
hipError_t k2_expand_count(hipStream_t s, uint32_t grid, const Item* in, const uint32_t* in_count, void* part,
                           uint32_t* out_count, const uint8_t* target_row, const float* pred = nullptr,
                           const float* thresh = nullptr);
hipError_t k2_expand_write(hipStream_t s, uint32_t grid, const Item* in, const uint32_t* in_count,
                           const void* part, bool interleave, Item* out, uint32_t out_cap,
                           const uint8_t* target_row, uint64_t* solve_nodes, uint64_t* exec_nodes, uint32_t* err,
                           const float* pred = nullptr, const float* thresh = nullptr);
// Split threshold of the learned stage, on the device: thresh = fixed if
// fixed > 0, else log2(max(1, mult * sum(2^pred) / count)), count read on the
// device; over = #(pred > thresh).
// The sum is formed in a fixed order (block_sums: kPredSumBlocks doubles of
// scratch), so equal inputs give equal thresholds. over must be zero on
// entry. Device counts are read up to max_count.
constexpr uint32_t kPredSumBlocks = 1024;
hipError_t k2_split_threshold(hipStream_t s, const float* pred, const uint32_t* count, uint32_t max_count,
                              float fixed, double mult, double* block_sums, double* sum, uint32_t* over,
                              float* thresh);
// Largest predicted first, dealt round-robin to nseg runs of equal size
// (nseg = the drain's queue segments).
hipError_t k2_order_by_prediction(hipStream_t s, const Item* in, const float* pred, const uint32_t* count,
                                  uint32_t max_count, uint32_t nseg, uint32_t* hist, Item* out);
hipError_t k2_drain(hipStream_t s, uint32_t blocks, uint32_t threads, const K2DrainParams& p);
// Queue items one drain wave takes per atomic on the queue head (at most);
// a budget round can hand back up to this many unstarted items per wave.
constexpr uint32_t kGrabItems = 128;
size_t k2_lds_bytes(uint32_t stack_rows, uint32_t threads);
int k2_max_blocks_per_cu(uint32_t threads, size_t lds_bytes, int* out);
// Drain workgroup size for a stack of stack_rows rows: the largest of
// 64/128/256 lanes that keeps all but at most one of the waves per CU that
// one-wave workgroups get. Measured: at 16 waves per CU, 256-lane workgroups
// beat one-wave workgroups at 17 by 1-3% (N=17, N=18, a 512-solve batch); at
// N=19's ten stack rows they would keep 12 of 15 and were 8% slower.
int k2_auto_threads(uint32_t stack_rows, uint32_t lds_per_block, uint32_t* threads);

// ---- subtree-size predictor (mlp.hip) -------------------------------------
// count samples, or, with d_count, the device count of them (count is then
// an upper bound that sizes the launch)
int mlp_predict_device(nqm_mlp* mlp, const void* d_x, uint32_t count, float* d_y,
                       const uint32_t* d_count = nullptr);
// Changes whenever the model's weights change (process-wide counter, so a
// new model never repeats an old model's value).
uint64_t mlp_version(const nqm_mlp* mlp);
hipError_t mlp_features_items(hipStream_t s, const Item* items, const uint32_t* count, void* feats,
                              uint32_t max_count);

}  // namespace nqm
