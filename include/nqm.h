/*
 * nqm.h -- C ABI of the MI355X-native N-Queens solution counter.
 *
 * Plain C types only (pointers, sizes, status codes). One shared library,
 * libnqm.so, built for gfx950 (wave64). Every entry point returns an
 * NQM_* status and never exits the process.
 *
 * Threading: one nqm_ctx per host thread. Calls on one context are
 * serialized by the caller; contexts are independent.
 */
#ifndef NQM_H
#define NQM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define NQM_OK 0
#define NQM_EINVAL 1       /* bad argument: n outside 1..31, depth > n, NULL, ... */
#define NQM_EPREFIX 2      /* a prefix in the pool is not a valid placement */
#define NQM_ENOMEM 3       /* host or device allocation failed */
#define NQM_EDEVICE 4      /* HIP runtime error, or not a gfx950 wave64 device */
#define NQM_ERANGE 5       /* output capacity too small / request over a limit */
#define NQM_EIO 6          /* file could not be opened, read or written */
#define NQM_EFORMAT 7      /* pool file is malformed or has the wrong version */
#define NQM_EFLAGS 8       /* pool file was written with other breaking flags */
#define NQM_EUNSUPPORTED 9 /* valid request this build does not serve */

/* ---- flags ------------------------------------------------------------- */
/* Weight a prefix x2 when its row-0 column is < n/2 (integer division):
 * the count of a pool holding only row-0 columns < ceil(n/2) then equals the
 * full-board count. Reference: GPU_NQueens.cu:354-357, :437-439. */
#define NQM_MIRROR_WEIGHT 0x1u

typedef struct nqm_ctx nqm_ctx;

/* Per-call measurements. kernel_ms is the HIP-event time of the counting
 * kernel(s) on the context's stream, comparable to the reference's event
 * timer (GPU_NQueens.cu:118, :143). nodes = queen placements executed below
 * the prefixes (rows >= depth). */
typedef struct {
    uint64_t total;        /* mirror-weighted solution count */
    uint64_t nodes;        /* placements executed below the prefixes */
    double kernel_ms;      /* device time of the counting launch(es) */
    double wall_ms;        /* host wall time of the whole call */
    uint32_t launches;     /* expansion levels and drain rounds the call ran (a
                              graph launch counts one; a graph-mode call that
                              builds its graph also ran the planning levels) */
    uint32_t blocks;       /* workgroups of the (first) counting launch */
    double search_ms;      /* device time of the search launches alone (K1 count;
                              K2 drain rounds), without expansion or splitting */
    /* Budgeted drain rounds (nqm_solve_batch with a budget; the server), for
     * checking that rounds do not multiply the work: a round parks at most
     * one piece per lane plus one split-off bottom row per lane, and hands
     * back at most the queue items its waves held (waves x items per grab). */
    uint32_t rounds;          /* drain launches */
    uint32_t reserved;
    uint64_t lanes;           /* lanes of a drain launch (the largest) */
    uint64_t parked_max;      /* most pieces one round parked (<= 2 x lanes) */
    uint64_t handed_back_max; /* most unstarted items one round handed back */
    uint64_t hand_back_room;  /* what one round may hand back (<= this) */
} nqm_stats;

/* Create a context on HIP device `device` (its own stream, events and
 * device buffers). Fails with NQM_EDEVICE unless the device is wave64. */
int nqm_create(int device, nqm_ctx** out);
void nqm_destroy(nqm_ctx* ctx);
const char* nqm_strerror(int status);
/* Detail of the last failure on this context ("" if none). */
const char* nqm_last_error(const nqm_ctx* ctx);
/* Device facts used by the roofline: compute units, wave size, peak engine
 * clock in kHz, LDS bytes per workgroup and per CU. */
int nqm_device_info(const nqm_ctx* ctx, uint32_t* cus, uint32_t* wave_size,
                    uint32_t* clock_khz, uint32_t* lds_per_block, uint32_t* lds_per_cu);

/* ---- the counting path (drop-in boundary) ------------------------------- */

/* Replaces gpu_solver_driver (GPU_NQueens.cuh:13, GPU_NQueens.cu:56-175)
 * for one device: a pool of `count` prefixes (host memory, count x depth
 * column bytes, row-major; column < n, rows 0..depth-1) goes in, the total
 * solution count comes out. n and depth are runtime values (the reference
 * fixes N at compile time, deffinitions.cuh:38). Static assignment, one
 * prefix per wave64 lane, per-block u64 partials summed on the host.
 * count == 0 is NQM_EINVAL, as the reference refuses it (GPU_NQueens.cu:58). */
int nqm_count_pool(nqm_ctx* ctx, uint32_t n, uint32_t depth, const uint8_t* pool,
                   uint64_t count, uint32_t flags, uint64_t* total_out);

/* Same counting kernel over a pool already resident in device memory
 * (d_pool from nqm_dev_alloc). d_per_sols / d_per_nodes (device, optional)
 * receive each prefix's unweighted solution and node counts. */
int nqm_count_pool_device(nqm_ctx* ctx, uint32_t n, uint32_t depth, const uint8_t* d_pool,
                          uint64_t count, uint32_t flags, uint64_t* d_per_sols,
                          uint64_t* d_per_nodes, nqm_stats* stats);

/* The same count with no allocation of any kind (no hipMalloc/hipFree, no
 * memset, no host synchronization): no counterpart in the reference; it is
 * for callers that cannot allocate, such as code under HIP graph capture or
 * a tight serving loop. Its scratch is a fixed 16 MiB per device, so pools
 * of more than 2^26 = 67,108,864 prefixes are NQM_ERANGE. Results go to
 * d_result (device, 3 x u64, written in the context's stream order):
 * [0] = mirror-weighted solution count, [1] = nodes below the prefixes,
 * [2] = nonzero if a prefix was invalid (the synchronous calls return
 * NQM_EPREFIX instead). All calls on one device share the scratch: calls
 * that overlap on one device race, so the caller serializes them. */
int nqm_count_pool_device_static(nqm_ctx* ctx, uint32_t n, uint32_t depth, const uint8_t* d_pool,
                                 uint64_t count, uint32_t flags, uint64_t* d_result);

/* The reference's entry point with its own argument layout: `states` is an
 * array of the reference's nq_state_t for board size n (n_queens.cuh:100-113:
 * u64 diagonal, u64 antidiagonal, u32 queens_in_columns, char curr_row,
 * u8 queen_at_index[n], padded to 8 bytes), every state with
 * curr_row == locked_row_end. Mirror weighting is always applied, as the
 * reference's build does (deffinitions.cuh:30). config_cnt must be 1: this
 * build drives one device per call. */
typedef struct {
    unsigned int device_id;
    int async; /* accepted for layout compatibility; calls are synchronous */
} nqm_gpu_config;
int nqm_solver_driver(uint32_t n, const void* states, uint32_t state_cnt,
                      unsigned locked_row_end, const nqm_gpu_config* configs,
                      unsigned config_cnt, uint64_t* total_out);
/* sizeof(nq_state_t) of the reference for board size n. */
size_t nqm_state_size(uint32_t n);

/* ---- prefix-pool generation (on the device) -----------------------------
 * The reference's pool: every valid placement of rows 0..lock-1 whose row
 * `lock` still has a free cell, in lexicographic order; with
 * NQM_MIRROR_WEIGHT only row-0 columns < ceil(n/2) (the half pool the
 * reference's build always generates). Replaces gen_states_rowlock and its
 * threaded drivers (n_queens.cu:27-83, :235-315); byte-identical output.
 * Requires 2 <= n <= 31 and 1 <= lock < n. */
int nqm_pool_count(nqm_ctx* ctx, uint32_t n, uint32_t lock, uint32_t flags,
                   uint64_t* count, uint64_t* per_col /* n entries, optional */);
/* Host output: count x lock bytes; NQM_ERANGE (with *count set) if cap is short. */
int nqm_pool_generate(nqm_ctx* ctx, uint32_t n, uint32_t lock, uint32_t flags,
                      uint8_t* out, uint64_t cap, uint64_t* count);
/* Device output, allocated by the library; release with nqm_dev_free. */
int nqm_pool_generate_device(nqm_ctx* ctx, uint32_t n, uint32_t lock, uint32_t flags,
                             uint8_t** d_out, uint64_t* count);
/* The reference's lock-row search for a requested pool size (nq_generate_states,
 * n_queens.cu:128-219): the lock whose pool the reference would return, and
 * that pool's size. NQM_ERANGE where the reference would FAIL_IF. */
int nqm_choose_lock(nqm_ctx* ctx, uint32_t n, uint64_t how_many, uint32_t flags,
                    uint32_t* lock_out, uint64_t* count_out);

/* ---- pool files and the reference's state layout (host only) ------------
 * .nqsf, the reference's pool file (nq_utils.cu:127-259): u32 version (1),
 * u64 breaking-flag word, u8 N, u8 lock, u64 total, then per populated row
 * h: u8 h, u64 count, count x h column bytes. Flag word bit 0 = the
 * reference's experimental-optimisation build flag, bit 1 = mirror-halved
 * pool (nq_utils.cu:65-74); the reference's default build writes 3. */
#define NQM_NQSF_FLAG_EXPERIMENTAL 0x1ull
#define NQM_NQSF_FLAG_HALF 0x2ull
int nqm_nqsf_write(const char* path, uint32_t n, uint32_t lock, const uint8_t* pool,
                   uint64_t count, uint64_t flags);
/* Reads a pool whose prefixes all have `lock` rows. expect_n = 0 skips the
 * board-size check; expect_flags = UINT64_MAX skips the flag check (else a
 * mismatch is NQM_EFLAGS, as util_read_nq_states_from_stream returns -4).
 * *pool is malloc'ed by the library: release with nqm_free. */
int nqm_nqsf_read(const char* path, uint32_t expect_n, uint64_t expect_flags,
                  uint8_t** pool, uint64_t* count, uint32_t* n_out, uint32_t* lock_out,
                  uint64_t* flags_out);
void nqm_free(void* p);
/* Reference nq_state_t array (see nqm_solver_driver) <-> compact pool. */
int nqm_states_to_pool(uint32_t n, const void* states, uint64_t count, uint32_t lock,
                       uint8_t* pool_out);
int nqm_pool_to_states(uint32_t n, uint32_t lock, const uint8_t* pool, uint64_t count,
                       void* states_out);

/* ---- subtree-size predictor (learned splitting) --------------------------
 * No reference counterpart. A 12 -> 64 -> 64 -> 1 MLP (ReLU, FP16 weights,
 * FP32 accumulation on MFMA) estimating log2(1 + nodes below a prefix) from
 * relative next-row bitsets. Trained on the device by the library's own
 * forward/backward kernels (Adam on FP32 master weights). Features are
 * nqm_mlp_feature_width() halves per sample; parameters are
 * nqm_mlp_param_count() floats: W1[16][64], b1[64], W2[64][64], b2[64],
 * w3[64], b3. */
typedef struct nqm_mlp nqm_mlp;
uint32_t nqm_mlp_param_count(void);
uint32_t nqm_mlp_feature_width(void);
int nqm_mlp_create(nqm_ctx* ctx, uint64_t seed, nqm_mlp** out);
void nqm_mlp_destroy(nqm_mlp* mlp);
int nqm_mlp_set_params(nqm_mlp* mlp, const float* params);
int nqm_mlp_get_params(nqm_mlp* mlp, float* params);
/* Features of every prefix of a device pool into d_feats (device, count rows). */
int nqm_mlp_features_pool(nqm_ctx* ctx, uint32_t n, uint32_t depth, const uint8_t* d_pool,
                          uint64_t count, void* d_feats);
int nqm_mlp_predict(nqm_mlp* mlp, const void* d_feats, uint32_t count, float* d_out);
/* Mean-squared-error gradient over `count` samples (host output, param order). */
int nqm_mlp_gradient(nqm_mlp* mlp, const void* d_feats, const float* d_labels, uint32_t count,
                     float* grad_out, float* loss_out);
/* Minibatch Adam; a seeded split keeps the last val_frac of samples for
 * validation. Losses are mean squared errors of log2(1 + nodes). */
int nqm_mlp_train(nqm_mlp* mlp, const void* d_feats, const float* d_labels, uint32_t count,
                  uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float val_frac,
                  float* train_loss, float* val_loss);

/* ---- batched persistent counting ----------------------------------------
 * Many independent solves (any mix of n and pre-placed prefixes) in one
 * persistent device pass. No reference counterpart: the reference counts one
 * pool of one N per launch (GPU_NQueens.cu:128-138); its README names
 * restarts as future work (README.md:46-56). Counts are exactly those of the
 * static path. */
typedef struct {
    uint32_t n;          /* board size, 1..31 */
    uint32_t k;          /* queens pre-placed in rows 0..k-1 (0..n) */
    uint8_t prefix[32];  /* their columns */
} nqm_solve;

#define NQM_TUNE_NO_MIRROR 0x1u  /* do not halve empty boards by mirror symmetry */
#define NQM_TUNE_NO_DONATE 0x2u  /* no in-wave work donation (queue refill only) */
#define NQM_TUNE_GRAPH 0x4u      /* run expand (+ learned stage) + drain as one
                                    captured HIP graph; a call repeating the
                                    previous call's solves, tuning and model
                                    weights launches that graph again */
#define NQM_TUNE_NO_ORDER 0x8u   /* with a model: split, but keep the queue order */

struct nqm_mlp;

typedef struct {
    uint32_t depth;        /* rows placed before the queue (per solve, capped
                              at n-1 and raised to k); 0 = automatic */
    uint32_t waves_per_cu; /* target resident waves per CU; 0 = as many as fit */
    uint32_t threads;      /* workgroup size 64/128/256; 0 = automatic: the
                              largest that keeps all but at most one of the
                              waves per CU of one-wave workgroups (64 when
                              waves_per_cu is set) */
    uint32_t refill_idle;  /* idle lanes of a wave that trigger refill/donation; 0 = 12.
                              A single-launch drain (budget 0) also gives each busy lane
                              its next queue item, which it starts as soon as its piece
                              ends; a wave refills once 3 x refill_idle lanes (at most
                              56) have started theirs */
    uint64_t budget;       /* passes per wave before unfinished pieces are parked
                              and the drain relaunched (restart); 0 = one launch */
    uint32_t flags;        /* NQM_TUNE_* */
    uint32_t stack_rows;   /* LDS stack rows per lane (the expansion then goes deep
                              enough that no piece needs more); 0 = automatic: what
                              the deepest piece needs */
    /* Learned splitting: with a model, the expansion stops early, every item
     * whose predicted log2(1 + nodes) exceeds split_log2 is split one row
     * deeper, up to split_rounds times (0 = 1), the queue is ordered largest
     * predicted subtree first (dealt round-robin to the drain's queue
     * segments, so each starts with its share of the largest), and
     * (automatic depth) the expansion then
     * continues to the stack-depth goal keeping that order. split_log2 = 0
     * picks log2(8 x the mean predicted size of the level). */
    struct nqm_mlp* model;
    float split_log2;
    uint32_t split_rounds;
} nqm_tune;

/* Count every solve. counts[i] = solutions extending solve i's prefix;
 * nodes[i] (optional) = placements below the prefix, full-tree numbers
 * (mirror-halved work is weighted back). stats->nodes = placements
 * actually executed on the device. */
int nqm_solve_batch(nqm_ctx* ctx, const nqm_solve* solves, uint32_t nsolves,
                    const nqm_tune* tune, uint64_t* counts, uint64_t* nodes, nqm_stats* stats);

/* Autotuner: runs the batch under every combination of the candidate
 * expansion depths, waves per CU and LDS stack-row caps (zero-terminated
 * lists; an empty list means "automatic only"), `repeats` times each, and
 * returns the fastest tune (median kernel time). Every combination must give
 * the same counts, or the call fails with NQM_ERANGE. A combination whose
 * queue does not fit (a stack-row cap or depth that asks for more than 2^31
 * items or more memory than the device has) is not run: its row reads -1 ms.
 * The call fails only if no combination fits. results (optional, capacity
 * max_results) receive one row per combination: depth, waves_per_cu,
 * stack_rows, median ms. */
typedef struct {
    uint32_t depths[8];
    uint32_t waves_per_cu[8];
    uint32_t stack_rows[8];
    uint32_t repeats;      /* 0 = 3 */
    uint32_t base_flags;   /* NQM_TUNE_* applied to every trial */
} nqm_tune_space;
typedef struct {
    uint32_t depth, waves_per_cu, stack_rows, pad;
    double ms;
} nqm_tune_result;
int nqm_autotune(nqm_ctx* ctx, const nqm_solve* solves, uint32_t nsolves,
                 const nqm_tune_space* space, nqm_tune* best, nqm_tune_result* results,
                 uint32_t max_results, uint32_t* nresults);

/* ---- symmetry-reduced counting --------------------------------------------
 * No counterpart in the reference, which lists "transformation elimination"
 * as future work (README.md:38-45). Counts the classes of solutions of the
 * empty n x n board under its 8 symmetries (unique) and the solutions they
 * stand for (total, equal to the reference's count), by the published corner
 * / non-corner method: about 22% of the full tree's placements are made.
 * stats->nodes = those placements (not comparable with the other paths). */
int nqm_count_symmetric(nqm_ctx* ctx, uint32_t n, uint64_t* unique, uint64_t* total, nqm_stats* stats);

/* The same count by a chosen kernel: NQM_SYM_THREAD = one thread per host
 * item (any n), NQM_SYM_DRAIN = the persistent drain (10 <= n <= 20, else
 * NQM_EUNSUPPORTED), NQM_SYM_AUTO = the drain where it applies (what
 * nqm_count_symmetric does). Every path gives the same unique, total and
 * stats->nodes. stats->search_ms = the drain (or thread) kernel alone,
 * kernel_ms = from the first expansion to the last result. */
#define NQM_SYM_AUTO 0u
#define NQM_SYM_THREAD 1u
#define NQM_SYM_DRAIN 2u
int nqm_count_symmetric_path(nqm_ctx* ctx, uint32_t n, uint32_t path, uint64_t* unique, uint64_t* total,
                             nqm_stats* stats);

/* ---- full DoubleSweep counting ---------------------------------------------
 * The reference counts with DoubleSweep-light (unit propagation of the current
 * row, GPU_NQueens.cu:392-408) and names the full algorithm as future work
 * (README.md:6, :22, :40-41); its unbuilt attempt walks the rows centre-out
 * (GPU_NQueens.cu:207, :214-336). This search propagates over every open row
 * and column until no line is forced (a line with no free cell ends the
 * branch, one with a single free cell gets its queen) and branches on the
 * open row nearest the centre (m, m+1, m-1, ...; m = (n-1)/2), the root
 * mirror-halved. total = the reference's count; stats->nodes = this search's
 * placements (children and propagated queens), about 4x fewer than the light
 * search's at n = 15..18. 1 <= n <= 21 (else NQM_EUNSUPPORTED).
 * kernel_ms = from the first expansion level to the end of the drain,
 * search_ms = the drain alone. */
int nqm_count_fullsweep(nqm_ctx* ctx, uint32_t n, uint64_t* total, nqm_stats* stats);
/* The same with the number of states the device expansion stops at (the
 * first level holding at least target_states). 0 = automatic, what
 * nqm_count_fullsweep does: the first level that holds 2^19 states and
 * leaves each drain lane at most 9 stack rows (levels stay under 2^26
 * states). Any target gives the same total and stats->nodes. */
int nqm_count_fullsweep_ex(nqm_ctx* ctx, uint32_t n, uint32_t target_states, uint64_t* total,
                           nqm_stats* stats);

/* ---- serving: continuous batching of count requests ---------------------
 * No reference counterpart. Requests join a device-resident work queue
 * between budgeted drain rounds (new requests are served first); pieces of
 * work a round leaves unfinished are parked and resumed by the next round.
 * A request's result is ready the round its last item or parked piece ends.
 * Admission is by free counter slot and by queue memory. Each count equals
 * that of the request solved alone. */
typedef struct nqm_server nqm_server;
typedef struct {
    uint32_t max_active;   /* requests in flight (counter slots); 0 = 256 */
    uint32_t depth;        /* rows a new request is expanded to; 0 = automatic:
                              row 6, or deep enough for at most 11 stack rows
                              (n > 18), as far as the queue memory left allows */
    uint64_t budget;       /* passes per wave per round; 0 = 65536 */
    uint64_t queue_bytes;  /* device memory for queued work items (two buffers of
                              this size are allocated); 0 = 8 GiB */
    uint32_t flags;        /* NQM_TUNE_NO_MIRROR | NQM_TUNE_NO_DONATE */
    uint32_t reserved;
} nqm_server_cfg;
typedef struct {
    uint64_t tag;          /* the caller's id of the request */
    uint64_t count;        /* solutions extending the prefix */
    uint64_t nodes;        /* placements below the prefix (full-tree numbers) */
    uint32_t rounds;       /* drain rounds the request stayed in flight */
    uint32_t reserved;
    double admit_ms;       /* server clock at admission */
    double done_ms;        /* server clock when the result was ready */
} nqm_result;
int nqm_server_create(nqm_ctx* ctx, const nqm_server_cfg* cfg, nqm_server** out);
void nqm_server_destroy(nqm_server* srv);
/* Offer requests in order; *admitted = how many (a prefix of reqs) were taken. */
int nqm_server_submit(nqm_server* srv, const nqm_solve* reqs, const uint64_t* tags,
                      uint32_t nreq, uint32_t* admitted);
/* One drain round over everything queued; finished requests become pollable. */
int nqm_server_step(nqm_server* srv, uint32_t* active_out);
int nqm_server_poll(nqm_server* srv, nqm_result* out, uint32_t max, uint32_t* nout);
int nqm_server_info(const nqm_server* srv, uint32_t* active, uint64_t* queued_items,
                    uint64_t* rounds);
/* The server's rounds so far in the round fields of nqm_stats (rounds, lanes,
 * parked_max, handed_back_max, hand_back_room; the other fields are 0). */
int nqm_server_round_stats(const nqm_server* srv, nqm_stats* out);
/* Milliseconds since nqm_server_create (the clock of admit_ms / done_ms). */
double nqm_server_clock_ms(const nqm_server* srv);

/* ---- device memory helpers (so callers need no HIP headers) ------------ */
int nqm_dev_alloc(nqm_ctx* ctx, size_t bytes, void** d_ptr);
int nqm_dev_free(nqm_ctx* ctx, void* d_ptr);
int nqm_memcpy_h2d(nqm_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (bytes % 16 == 0): the
 * known-size reference for HBM byte counters. */
int nqm_dev_copy(nqm_ctx* ctx, void* d_dst, const void* d_src, size_t bytes, double* kernel_ms);
int nqm_memcpy_d2h(nqm_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* NQM_H */
