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
    uint32_t launches;     /* kernel launches the call made */
    uint32_t blocks;       /* workgroups of the (first) counting launch */
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
#define NQM_TUNE_GRAPH 0x4u      /* run expand + drain as one captured HIP graph */

typedef struct {
    uint32_t depth;        /* rows placed before the queue (per solve, capped
                              at n-1 and raised to k); 0 = automatic */
    uint32_t waves_per_cu; /* target resident waves per CU; 0 = as many as fit */
    uint32_t threads;      /* workgroup size 64/128/256; 0 = 256 */
    uint32_t refill_idle;  /* idle lanes of a wave that trigger refill/donation; 0 = 16 */
    uint64_t budget;       /* passes per wave before live work is spilled and the
                              queue relaunched (restart); 0 = drain in one launch */
    uint32_t flags;        /* NQM_TUNE_* */
    uint32_t stack_rows;   /* LDS stack rows per lane; 0 = what the deepest solve needs */
} nqm_tune;

/* Count every solve. counts[i] = solutions extending solve i's prefix;
 * nodes[i] (optional) = placements below the prefix, full-tree numbers
 * (mirror-halved work is weighted back). stats->nodes = placements
 * actually executed on the device. */
int nqm_solve_batch(nqm_ctx* ctx, const nqm_solve* solves, uint32_t nsolves,
                    const nqm_tune* tune, uint64_t* counts, uint64_t* nodes, nqm_stats* stats);

/* ---- device memory helpers (so callers need no HIP headers) ------------ */
int nqm_dev_alloc(nqm_ctx* ctx, size_t bytes, void** d_ptr);
int nqm_dev_free(nqm_ctx* ctx, void* d_ptr);
int nqm_memcpy_h2d(nqm_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
int nqm_memcpy_d2h(nqm_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* NQM_H */
