/*
 * nq_fullsweep.c -- CPU oracle for full DoubleSweep counting.
 *
 * TEST INFRASTRUCTURE ONLY (see nq_oracle.h): the checker of the device's
 * full-sweep count, never linked by the product.
 *
 * The reference implements only the light variant (unit propagation of the
 * current row, GPU_NQueens.cu:392-408) and names the full algorithm as future
 * work (README.md:6, :22, :40-41, after the OKlibrary's DoubleSweep). Its one
 * attempt, the unbuilt kern_doitall_v2_regld_full_ds (GPU_NQueens.cu:214-336),
 * walks the rows centre-out (CALC_CURRENT_ROW, GPU_NQueens.cu:207: rows m,
 * m+1, m-1, m+2, ...). The OKlibrary source is not part of the reference, so
 * the search is defined here, from that description:
 *
 *   propagate(S): repeat { row sweep: every open row, ascending (the open rows
 *     when the sweep starts); no free cell: the board is dead; one free cell:
 *     a queen goes there (a placement). Column sweep: the same over the open
 *     columns, a column's free cells being its cells in open rows. } until a
 *     row sweep and the column sweep after it place nothing.
 *   node(S, w): propagate(S); dead: nothing; all rows placed: w solutions;
 *     otherwise branch on the open row nearest the centre row m = (n-1)/2 in
 *     the order m, m+1, m-1, m+2, ..., its free columns ascending (each child
 *     a placement), node(child, w).
 *   root: the empty board branches on row m; columns c < n/2 weigh 2 and the
 *     middle column of odd n weighs 1 (the mirror image of a left-half
 *     subtree holds as many solutions); the right half is not searched.
 *
 * Counts: solutions (weighted) and nodes = the placements the halved search
 * executes (children and propagated queens; unweighted).
 *
 * Pinning (tests/test_oracle.py): solutions equal the known totals of
 * test.cu:18; solutions and nodes equal an independent restatement on an
 * explicit board (tests/fullsweep_ref.py) for small n; nodes never exceed the
 * light search's; a subtree count from any state equals the sum over its
 * children (the recursion), and the multi-threaded count equals the
 * single-threaded one.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>

#include "nq_oracle.h"

typedef struct {
    uint32_t P, C;    /* rows / columns holding a queen */
    uint64_t D1, D2;  /* diagonals: bit r+c, bit c-r+n-1 */
} fs_state;

typedef struct {
    unsigned n;
    uint32_t M;
    uint64_t sols, nodes;
} fs_ctx;

static uint32_t fs_row_free(const fs_ctx* x, const fs_state* s, unsigned r) {
    return ~(s->C | (uint32_t)(s->D1 >> r) | (uint32_t)(s->D2 >> (x->n - 1u - r))) & x->M;
}

/* Free cells of column c as a row mask (cells of open rows only). */
static uint32_t fs_col_free(const fs_ctx* x, const fs_state* s, unsigned c) {
    uint32_t m = 0;
    for (unsigned r = 0; r < x->n; ++r) {
        const int blocked = ((s->P >> r) & 1u) || ((s->D1 >> (r + c)) & 1u) ||
                            ((s->D2 >> (c + x->n - 1u - r)) & 1u);
        if (!blocked) m |= 1u << r;
    }
    return m;
}

static void fs_place(const fs_ctx* x, fs_state* s, unsigned r, unsigned c) {
    s->P |= 1u << r;
    s->C |= 1u << c;
    s->D1 |= 1ull << (r + c);
    s->D2 |= 1ull << (c + x->n - 1u - r);
}

static unsigned fs_ctz(uint32_t v) { return (unsigned)__builtin_ctz(v); }

/* Returns 0 when the board is dead, 1 when stable; placements counted. */
static int fs_propagate(fs_ctx* x, fs_state* s) {
    for (;;) {
        int placed = 0;
        for (uint32_t o = ~s->P & x->M; o; o &= o - 1u) {
            const unsigned r = fs_ctz(o);
            const uint32_t f = fs_row_free(x, s, r);
            if (!f) return 0;
            if (!(f & (f - 1u))) {
                fs_place(x, s, r, fs_ctz(f));
                ++x->nodes;
                placed = 1;
            }
        }
        for (uint32_t o = ~s->C & x->M; o; o &= o - 1u) {
            const unsigned c = fs_ctz(o);
            const uint32_t g = fs_col_free(x, s, c);
            if (!g) return 0;
            if (!(g & (g - 1u))) {
                fs_place(x, s, fs_ctz(g), c);
                ++x->nodes;
                placed = 1;
            }
        }
        if (!placed) return 1;
    }
}

/* The open row nearest the centre row, the row below first on a tie. */
static unsigned fs_branch_row(const fs_ctx* x, uint32_t P) {
    const int m = (int)(x->n - 1u) / 2;
    for (int d = 0; d < (int)x->n; ++d) {
        const int lo = m + d, hi = m - d;
        if (lo < (int)x->n && !((P >> lo) & 1u)) return (unsigned)lo;
        if (d && hi >= 0 && !((P >> hi) & 1u)) return (unsigned)hi;
    }
    return 0u;
}

static void fs_node(fs_ctx* x, fs_state s, uint64_t w) {
    if (!fs_propagate(x, &s)) return;
    if (s.P == x->M) {
        x->sols += w;
        return;
    }
    const unsigned r = fs_branch_row(x, s.P);
    for (uint32_t f = fs_row_free(x, &s, r); f; f &= f - 1u) {
        fs_state t = s;
        fs_place(x, &t, r, fs_ctz(f));
        ++x->nodes;
        fs_node(x, t, w);
    }
}

int orc_fullsweep_subtree(unsigned n, uint32_t P, uint32_t C, uint64_t D1, uint64_t D2,
                          uint64_t* sols, uint64_t* nodes) {
    if (n < 1 || n > 31) return -1;
    fs_ctx x = {n, (uint32_t)((1ull << n) - 1ull), 0, 0};
    fs_state s = {P, C, D1, D2};
    fs_node(&x, s, 1u);
    if (sols) *sols = x.sols;
    if (nodes) *nodes = x.nodes;
    return 0;
}

int orc_fullsweep_children(unsigned n, uint32_t P, uint32_t C, uint64_t D1, uint64_t D2,
                           uint64_t* out, unsigned cap, uint64_t* prop_nodes) {
    if (n < 1 || n > 31) return -1;
    fs_ctx x = {n, (uint32_t)((1ull << n) - 1ull), 0, 0};
    fs_state s = {P, C, D1, D2};
    int k = 0;
    if (fs_propagate(&x, &s) && s.P != x.M) {
        const unsigned r = fs_branch_row(&x, s.P);
        for (uint32_t f = fs_row_free(&x, &s, r); f; f &= f - 1u) {
            fs_state t = s;
            fs_place(&x, &t, r, fs_ctz(f));
            if ((unsigned)k < cap) {
                out[4 * k + 0] = t.P;
                out[4 * k + 1] = t.C;
                out[4 * k + 2] = t.D1;
                out[4 * k + 3] = t.D2;
            }
            ++k;
        }
    }
    if (prop_nodes) *prop_nodes = x.nodes;
    return k;
}

/* The halved search split for threads: the root's children (row m, left-half
 * columns weighing 2, the middle column of odd n weighing 1) are propagated
 * and branched here, and the threads take their children from a shared
 * index. */
typedef struct {
    fs_state s;
    uint64_t w;
} fs_job;

typedef struct {
    unsigned n;
    fs_job* jobs;
    size_t njobs;
    size_t next;  /* shared, under lock */
    pthread_mutex_t lock;
    uint64_t sols, nodes;
} fs_pool;

static void* fs_worker(void* p) {
    fs_pool* q = (fs_pool*)p;
    fs_ctx x = {q->n, (uint32_t)((1ull << q->n) - 1ull), 0, 0};
    for (;;) {
        pthread_mutex_lock(&q->lock);
        const size_t i = q->next++;
        pthread_mutex_unlock(&q->lock);
        if (i >= q->njobs) break;
        fs_node(&x, q->jobs[i].s, q->jobs[i].w);
    }
    pthread_mutex_lock(&q->lock);
    q->sols += x.sols;
    q->nodes += x.nodes;
    pthread_mutex_unlock(&q->lock);
    return NULL;
}

int orc_fullsweep_count(unsigned n, unsigned threads, uint64_t* sols, uint64_t* nodes) {
    if (n < 1 || n > 31) return -1;
    fs_ctx x = {n, (uint32_t)((1ull << n) - 1ull), 0, 0};
    const unsigned m = (n - 1u) / 2u;
    fs_job* jobs = (fs_job*)malloc(sizeof(fs_job) * (size_t)n * n);
    if (!jobs) return -1;
    size_t nj = 0;
    for (unsigned c = 0; c < n / 2u + (n & 1u); ++c) {
        const uint64_t w = (c < n / 2u) ? 2u : 1u;
        fs_state s = {0, 0, 0, 0};
        fs_place(&x, &s, m, c);
        ++x.nodes;
        if (!fs_propagate(&x, &s)) continue;
        if (s.P == x.M) {
            x.sols += w;
            continue;
        }
        const unsigned r = fs_branch_row(&x, s.P);
        for (uint32_t f = fs_row_free(&x, &s, r); f; f &= f - 1u) {
            fs_state t = s;
            fs_place(&x, &t, r, fs_ctz(f));
            ++x.nodes;
            jobs[nj].s = t;
            jobs[nj].w = w;
            ++nj;
        }
    }
    fs_pool q;
    q.n = n;
    q.jobs = jobs;
    q.njobs = nj;
    q.next = 0;
    q.sols = x.sols;
    q.nodes = x.nodes;
    pthread_mutex_init(&q.lock, NULL);
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    pthread_t tid[64];
    unsigned started = 0;
    for (unsigned i = 1; i < threads; ++i)
        if (pthread_create(&tid[started], NULL, fs_worker, &q) == 0) ++started;
    fs_worker(&q);
    for (unsigned i = 0; i < started; ++i) pthread_join(tid[i], NULL);
    pthread_mutex_destroy(&q.lock);
    free(jobs);
    if (sols) *sols = q.sols;
    if (nodes) *nodes = q.nodes;
    return 0;
}
