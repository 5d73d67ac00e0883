/*
 * nq_oracle.c -- CPU oracle for the N-Queens counting path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product). Each function
 * cites the reference file:line it restates. Plain C, gcc -O2.
 */
#include "nq_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define ORC_UNSET 0xFFu /* n_queens.cuh:12 UNSET_QUEEN_INDEX */
#define ORC_MAXN 32

typedef struct {
    uint64_t diag, anti;   /* diagonals.cuh:8-10: bit c+r and bit c+n-r */
    uint32_t cols;         /* n_queens.cuh:108 */
    int row;               /* n_queens.cuh:110 curr_row */
    uint8_t q[ORC_MAXN];   /* n_queens.cuh:112 queen_at_index */
} orc_state;

static inline uint32_t orc_mask(unsigned n) { return n >= 32 ? 0xFFFFFFFFu : ((1u << n) - 1u); }

/* diagonals.cuh:17-20 dad_extract_explicit, n_queens.cuh:14 NQ_FREE_COLS_AT */
static inline uint32_t orc_free(const orc_state* s, unsigned n, int row) {
    uint32_t att = (uint32_t)((s->diag >> row) | (s->anti >> (n - (unsigned)row)));
    return ~(s->cols | att) & orc_mask(n);
}

/* n_queens.cuh:125-132 place_queen_at with diagonals.cuh:26-29 dad_add */
static inline void orc_place(orc_state* s, unsigned n, int row, unsigned col) {
    s->cols |= 1u << col;
    s->q[row] = (uint8_t)col;
    s->diag |= ((uint64_t)1u << col) << row;
    s->anti |= ((uint64_t)1u << col) << (n - (unsigned)row);
}

/* n_queens.cuh:134-141 remove_queen_at with diagonals.cuh:31-34 dad_remove */
static inline void orc_remove(orc_state* s, unsigned n, int row, unsigned col) {
    s->cols &= ~(1u << col);
    s->q[row] = ORC_UNSET;
    s->diag &= ~(((uint64_t)1u << col) << row);
    s->anti &= ~(((uint64_t)1u << col) << (n - (unsigned)row));
}

static inline unsigned orc_ctz(uint32_t x) { return (unsigned)__builtin_ctz(x); }
static inline unsigned orc_popc(uint32_t x) { return (unsigned)__builtin_popcount(x); }

/* Build the state the pool generator hands the kernel: rows 0..depth-1
 * placed, curr_row = depth (n_queens.cu:51-52; nq_utils.cu:244-251 rebuilds
 * states from column bytes the same way). Returns -1 on a conflict. */
static int orc_state_from_prefix(orc_state* s, unsigned n, unsigned depth, const uint8_t* cols) {
    memset(s, 0, sizeof(*s));
    memset(s->q, ORC_UNSET, sizeof(s->q));
    for (unsigned r = 0; r < depth; ++r) {
        unsigned c = cols[r];
        if (c >= n) return -1;
        if (!(orc_free(s, n, (int)r) & (1u << c))) return -1;
        orc_place(s, n, (int)r, c);
    }
    s->row = (int)depth;
    return 0;
}

/* The per-thread body of kern_doitall_v2_regld, GPU_NQueens.cu:344-436,
 * for one lane: the outer do-while ends when row < lock (the lane's vote in
 * the ballot at :359-360); each pass does one ADVANCE placement (:363-389),
 * then unit-row propagation (:392-408), then adds (cols == mask) (:410-435).
 * Node counting (one per placement) has no counterpart in the reference. */
int orc_count_prefix(unsigned n, unsigned depth, const uint8_t* cols,
                     uint64_t* sols, uint64_t* nodes) {
    if (n == 0 || n > 31 || depth > n) return -1;
    orc_state s;
    if (orc_state_from_prefix(&s, n, depth, cols) != 0) return -1;
    const uint32_t m = orc_mask(n);
    const int lock = (int)depth;
    uint64_t cnt = 0, nd = 0;
    if (depth == n) { /* a full board: one solution, nothing below */
        *sols = 1; *nodes = 0;
        return 0;
    }
    for (;;) {
        if (s.row < lock) break;                       /* :359-360 */
        while (s.row >= lock) {                        /* :363 */
            const unsigned qi = s.q[s.row];
            uint32_t f = orc_free(&s, n, s.row);        /* :367 */
            if (qi != ORC_UNSET) {                     /* :369-373 */
                f &= (m << (qi + 1)) & m;
                orc_remove(&s, n, s.row, qi);
            }
            if (!f) {
                --s.row;                               /* :375-376 */
            } else {
                orc_place(&s, n, s.row, orc_ctz(f));   /* :381-382 */
                ++nd;
                if (s.row < (int)n - 1) ++s.row;       /* :385-386 */
                break;
            }
        }
        /* :393 (the row >= 0 guard only matters for depth 0, which the
         * reference never runs: it assumes lock >= 1, GPU_NQueens.cu:355) */
        while (s.row >= 0 && s.q[s.row] == ORC_UNSET) {
            const uint32_t f = orc_free(&s, n, s.row);
            if (orc_popc(f) == 1) {                    /* :395-396 */
                orc_place(&s, n, s.row, orc_ctz(f));   /* :399-403 */
                ++nd;
                if (s.row < (int)n - 1) ++s.row;       /* :404 */
            } else {
                break;
            }
        }
        cnt += (s.cols == m);                          /* :434 */
    }
    *sols = cnt;
    *nodes = nd;
    return 0;
}

typedef struct {
    unsigned n, depth;
    const uint8_t* pool;
    uint64_t count;
    int mirror;
    uint64_t* per_sols;
    uint64_t* per_nodes;
    uint64_t next;           /* dynamic scheduling cursor */
    pthread_mutex_t lock;
    uint64_t total, nodes;
    int bad;
} orc_pool_job;

#define ORC_CHUNK 64

static void* orc_pool_worker(void* arg) {
    orc_pool_job* j = (orc_pool_job*)arg;
    uint64_t tot = 0, nd = 0;
    int bad = 0;
    for (;;) {
        uint64_t b = __atomic_fetch_add(&j->next, ORC_CHUNK, __ATOMIC_RELAXED);
        if (b >= j->count) break;
        uint64_t e = b + ORC_CHUNK < j->count ? b + ORC_CHUNK : j->count;
        for (uint64_t i = b; i < e; ++i) {
            const uint8_t* p = j->pool + i * j->depth;
            uint64_t s = 0, k = 0;
            if (orc_count_prefix(j->n, j->depth, p, &s, &k) != 0) { bad = 1; continue; }
            if (j->per_sols) j->per_sols[i] = s;
            if (j->per_nodes) j->per_nodes[i] = k;
            /* GPU_NQueens.cu:356 and :437-439: x2 when row-0 column < n>>1 */
            if (j->mirror && j->depth > 0 && p[0] < (j->n >> 1)) s <<= 1;
            tot += s;
            nd += k;
        }
    }
    pthread_mutex_lock(&j->lock);
    j->total += tot;
    j->nodes += nd;
    j->bad |= bad;
    pthread_mutex_unlock(&j->lock);
    return NULL;
}

int orc_count_pool(unsigned n, unsigned depth, const uint8_t* pool,
                   uint64_t count, int mirror, unsigned threads,
                   uint64_t* total, uint64_t* nodes,
                   uint64_t* per_sols, uint64_t* per_nodes) {
    orc_pool_job j;
    memset(&j, 0, sizeof(j));
    j.n = n; j.depth = depth; j.pool = pool; j.count = count; j.mirror = mirror;
    j.per_sols = per_sols; j.per_nodes = per_nodes;
    pthread_mutex_init(&j.lock, NULL);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    unsigned started = 0;
    for (unsigned t = 1; t < threads; ++t)
        if (pthread_create(&tid[started], NULL, orc_pool_worker, &j) == 0) ++started;
    orc_pool_worker(&j);
    for (unsigned t = 0; t < started; ++t) pthread_join(tid[t], NULL);
    pthread_mutex_destroy(&j.lock);
    if (total) *total = j.total;
    if (nodes) *nodes = j.nodes;
    return j.bad ? -1 : 0;
}

/* locked_advance_nq_state, n_queens.cu:423-445: one placement in rows
 * [top, ...], never stepping past row `lock`; false when the search leaves
 * the rows above `top`. */
static int orc_locked_advance(orc_state* s, unsigned n, int top, int lock) {
    const uint32_t m = orc_mask(n);
    while (s->row >= top) {
        const unsigned qi = s->q[s->row];
        uint32_t f = orc_free(s, n, s->row);
        if (qi != ORC_UNSET) {
            f &= (m << (qi + 1)) & m;
            orc_remove(s, n, s->row, qi);
        }
        if (f == 0) {
            --s->row;
        } else {
            orc_place(s, n, s->row, orc_ctz(f));
            if (s->row < lock) s->row++;
            return 1;
        }
    }
    return 0;
}

/* gen_states_rowlock, n_queens.cu:27-83, for the master state with a queen
 * at (0, c0) and curr_row 1 (n_queens.cu:114-116). Calls emit() for every
 * state whose rows 0..lock-1 are placed and whose row `lock` still has a free
 * cell (:47-49). lock >= 2 here; lock == 1 is handled by the caller (the
 * reference's lock==1 path emits depth-2 states, a quirk not reproduced). */
typedef int (*orc_emit_fn)(void* ctx, const orc_state* s);

static int orc_gen_rowlock(unsigned n, unsigned c0, unsigned lock, orc_emit_fn emit, void* ctx) {
    orc_state s;
    memset(&s, 0, sizeof(s));
    memset(s.q, ORC_UNSET, sizeof(s.q));
    orc_place(&s, n, 0, c0);
    s.row = 1;
    if (lock == 1) {
        if (orc_free(&s, n, 1)) return emit(ctx, &s);
        return 0;
    }
    for (;;) {
        int adv;
        do {
            adv = orc_locked_advance(&s, n, 1, (int)lock - 1);
        } while (adv && (!orc_free(&s, n, s.row + 1) || s.q[s.row] == ORC_UNSET));
        if (!adv) break;
        if (emit(ctx, &s) != 0) return -1;
    }
    return 0;
}

typedef struct { uint64_t cnt; } orc_count_ctx;
static int orc_emit_count(void* ctx, const orc_state* s) { (void)s; ((orc_count_ctx*)ctx)->cnt++; return 0; }

/* The row-0 column range: ceil(n/2) with mirror halving (n_queens.cu:130-134
 * GEN_STATES_TCNT), else n. */
static unsigned orc_row0_span(unsigned n, int mirror) { return mirror ? (n + 1) / 2 : n; }

uint64_t orc_pool_count(unsigned n, unsigned lock, int mirror, uint64_t* per_col) {
    if (n == 0 || n > 31 || lock == 0 || lock >= n) return 0;
    uint64_t tot = 0;
    const unsigned span = orc_row0_span(n, mirror);
    for (unsigned c = 0; c < span; ++c) {
        orc_count_ctx cc = {0};
        orc_gen_rowlock(n, c, lock, orc_emit_count, &cc);
        if (per_col) per_col[c] = cc.cnt;
        tot += cc.cnt;
    }
    return tot;
}

typedef struct { uint8_t* out; uint64_t cap, cnt; unsigned lock; } orc_write_ctx;
static int orc_emit_write(void* ctx, const orc_state* s) {
    orc_write_ctx* w = (orc_write_ctx*)ctx;
    if (w->cnt >= w->cap) return -1;
    memcpy(w->out + w->cnt * w->lock, s->q, w->lock);
    w->cnt++;
    return 0;
}

uint64_t orc_pool_generate(unsigned n, unsigned lock, int mirror, uint8_t* out, uint64_t cap) {
    if (n == 0 || n > 31 || lock == 0 || lock >= n) return 0;
    orc_write_ctx w = {out, cap, 0, lock};
    const unsigned span = orc_row0_span(n, mirror);
    for (unsigned c = 0; c < span; ++c) /* concatenated in column order, n_queens.cu:208-213 */
        if (orc_gen_rowlock(n, c, lock, orc_emit_write, &w) != 0) return UINT64_MAX;
    return w.cnt;
}

/* gen_states_threaded, n_queens.cu:128-219, replayed on pool sizes only:
 * start at floor(ln k / ln n) (:137-139), keep the previous pool while the
 * next is not smaller (downslope bound, :188-189), lock <= n and the pool is
 * <= k*1.2 (:135, :191). A per-column count reaching the cap is the
 * generator's ERANGE (:76-80), which ends the search (:177-180). */
int orc_lock_depth(unsigned n, uint64_t k, int mirror, unsigned* lock_out, uint64_t* count_out) {
    if (n < 2 || n > 31 || k == 0) return -1;
    const uint64_t cap = (uint64_t)((double)k * (1.0 + 0.2));
    unsigned lock = (unsigned)(log((double)k) / log((double)n));
    if (lock >= n) lock = n - 1;
    if (lock == 0) return -1;
    int have_actual = 0;
    uint64_t actual = 0, tmp = 0;
    unsigned actual_lock = 0, tmp_lock = 0;
    int have_tmp = 0;
    const unsigned span = orc_row0_span(n, mirror);
    for (;;) {
        actual = tmp; actual_lock = tmp_lock; have_actual = have_tmp;
        have_tmp = 0; tmp = 0;
        if (lock >= n) break; /* pools exist for lock <= n-1 here */
        uint64_t total = 0;
        int err = 0;
        for (unsigned c = 0; c < span; ++c) {
            orc_count_ctx cc = {0};
            orc_gen_rowlock(n, c, lock, orc_emit_count, &cc);
            if (cc.cnt >= cap) { err = 1; break; }
            total += cc.cnt;
        }
        if (err) break;
        tmp = total; tmp_lock = lock; have_tmp = 1;
        lock++;
        if (!(tmp >= actual && lock <= n && tmp <= cap)) break;
    }
    if (!have_actual) return -1;
    *lock_out = actual_lock;
    *count_out = actual;
    return 0;
}

static void orc_levels_rec(unsigned n, unsigned maxk, unsigned row, uint32_t cols, uint32_t ld,
                           uint32_t rd, uint64_t* lv) {
    const uint32_t m = orc_mask(n);
    uint32_t f = ~(cols | ld | rd) & m;
    while (f) {
        const uint32_t b = f & (0u - f);
        f ^= b;
        lv[row + 1]++;
        if (row + 1 < maxk)
            orc_levels_rec(n, maxk, row + 1, cols | b, ((ld | b) << 1) & m, (rd | b) >> 1, lv);
    }
}

void orc_level_counts(unsigned n, unsigned max_k, uint64_t* levels) {
    if (max_k > n) max_k = n;
    for (unsigned k = 0; k <= max_k; ++k) levels[k] = 0;
    levels[0] = 1;
    if (n == 0 || n > 31 || max_k == 0) return;
    orc_levels_rec(n, max_k, 0, 0, 0, 0, levels);
}

static uint8_t* orc_put(uint8_t* p, const void* v, size_t len) {
    memcpy(p, v, len);
    return p + len;
}

/* util_write_nq_states_to_stream, nq_utils.cu:145-193, with
 * COMPILATION_PARAMETER_FLAGS_IN_STATE_FILES defined (deffinitions.cuh:83):
 * u32 version (=1, nq_utils.cu:13), u64 flags, u8 n, u8 lock, u64 total, then
 * one segment per populated row: u8 row, u64 count, count*row column bytes.
 * All prefixes of a pool share one row here (= lock). */
uint64_t orc_nqsf_write(unsigned n, unsigned lock, const uint8_t* pool,
                        uint64_t count, uint64_t flags, uint8_t* buf, uint64_t cap) {
    const uint64_t need = 4 + 8 + 1 + 1 + 8 + (count ? (1 + 8 + count * lock) : 0);
    if (need > cap) return 0;
    uint8_t* p = buf;
    const uint32_t version = 1;
    const uint8_t en = (uint8_t)n, lk = (uint8_t)lock;
    p = orc_put(p, &version, 4);
    p = orc_put(p, &flags, 8);
    p = orc_put(p, &en, 1);
    p = orc_put(p, &lk, 1);
    p = orc_put(p, &count, 8);
    if (count) {
        p = orc_put(p, &lk, 1);
        p = orc_put(p, &count, 8);
        p = orc_put(p, pool, count * lock);
    }
    return (uint64_t)(p - buf);
}
