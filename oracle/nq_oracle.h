/*
 * nq_oracle.h -- CPU oracle for the N-Queens counting path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product (gpu-n-queens_amd/) never links or calls it.
 *
 * It restates, in plain C, the algorithm of the reference fillpant/GPU-N-Queens:
 *   - per-prefix counting loop (advance + unit-row propagation + count),
 *     GPU_NQueens.cu:338-452, with the absolute diagonal encoding of
 *     diagonals.cuh:17-34 and the place/remove steps of n_queens.cuh:125-141;
 *   - mirror weighting, GPU_NQueens.cu:354-357 and :437-439;
 *   - prefix-pool generation, n_queens.cu:27-83 (gen_states_rowlock) driven
 *     per row-0 column as in n_queens.cu:111-124 and concatenated in column
 *     order as in n_queens.cu:203-213;
 *   - the lock-depth heuristic, n_queens.cu:128-219;
 *   - the .nqsf pool file layout, nq_utils.cu:127-259.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   orc_count_prefix / orc_count_pool : known-answer table test.cu:18, brute
 *       force over permutations (N<=9), mirror closed form, node recursion.
 *   orc_pool_generate / orc_pool_count : the 17 (N, lock, count) triples in
 *       the pool file names at test.cu:88-104, brute force for small N.
 *   orc_lock_depth : replay of n_queens.cu:135-192 by hand for N=16/18 and
 *       the derived pool sizes; no reference run exists (parity unpinned
 *       beyond these replays).
 *   orc_nqsf_* : byte layout from the format comment nq_utils.cu:127-144;
 *       no reference-written file exists (parity unpinned beyond layout).
 *   orc_fullsweep_* (nq_fullsweep.c): solutions by test.cu:18; solutions and
 *       nodes by an independent explicit-board restatement for small n; the
 *       algorithm itself is defined in nq_fullsweep.c from the reference's
 *       description (the OKlibrary source is not part of the reference).
 */
#ifndef NQ_ORACLE_H
#define NQ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Count solutions and nodes (queen placements in rows >= depth) below one
 * prefix. cols[r] is the column of the queen in row r, r < depth.
 * Returns 0, or -1 if the prefix is not a valid partial placement. */
int orc_count_prefix(unsigned n, unsigned depth, const uint8_t* cols,
                     uint64_t* sols, uint64_t* nodes);

/* Count a whole pool (count x depth bytes, row-major). If mirror != 0 a
 * prefix whose row-0 column is < n/2 (integer division) counts twice, as the
 * reference kernel does. Work is split over `threads` pthreads with dynamic
 * scheduling. per_sols / per_nodes (optional) receive unweighted per-prefix
 * results. Returns 0 or -1 (invalid prefix found). */
int orc_count_pool(unsigned n, unsigned depth, const uint8_t* pool,
                   uint64_t count, int mirror, unsigned threads,
                   uint64_t* total, uint64_t* nodes,
                   uint64_t* per_sols, uint64_t* per_nodes);

/* Number of prefixes of the reference's pool at lock row `lock`
 * (rows 0..lock-1 placed, row `lock` keeps >= 1 free cell). With mirror,
 * only row-0 columns < ceil(n/2) are used. per_col (optional, n entries)
 * receives the count per row-0 column. */
uint64_t orc_pool_count(unsigned n, unsigned lock, int mirror, uint64_t* per_col);

/* Write the pool (lock bytes per prefix) into out (capacity cap prefixes).
 * Returns the number of prefixes, or UINT64_MAX if cap is too small. */
uint64_t orc_pool_generate(unsigned n, unsigned lock, int mirror,
                           uint8_t* out, uint64_t cap);

/* The reference's lock-depth search for a requested pool size k.
 * Returns 0 and writes the chosen lock row and pool size, or -1 when the
 * reference would fail (FAIL_IF at n_queens.cu:139 or :196). */
int orc_lock_depth(unsigned n, uint64_t k, int mirror,
                   unsigned* lock_out, uint64_t* count_out);

/* Number of valid partial placements of exactly k queens in rows 0..k-1,
 * for k = 0..max_k (levels[0] = 1; max_k <= n). With max_k = n, summing
 * levels[1..n] gives the full-tree node count. Plain recursive DFS using the
 * relative bitset encoding, independent of orc_count_prefix. */
void orc_level_counts(unsigned n, unsigned max_k, uint64_t* levels);

/* Symmetry-reduced counting (nq_symmetry.c; the reference's future work,
 * README.md:38-45): solution classes under the 8 board symmetries (unique),
 * the solutions they stand for (total) and the placements the reduced search
 * makes (nodes). Pinned by orc_sym_brute and OEIS A002562. */
int orc_sym_count(unsigned n, uint64_t* unique, uint64_t* total, uint64_t* nodes);
/* Every solution and its 8 images; a class is counted at its smallest member
 * (row by row). n <= 16. Returns -2 if the classes' orbits do not add up. */
int orc_sym_brute(unsigned n, uint64_t* unique, uint64_t* total);

/* Full DoubleSweep counting (nq_fullsweep.c, which defines the search):
 * solutions and placements of the mirror-halved search of an n x n board,
 * split over up to `threads` threads by the root's columns. */
int orc_fullsweep_count(unsigned n, unsigned threads, uint64_t* sols, uint64_t* nodes);
/* Solutions and placements below one state (queens already placed: rows P,
 * columns C, diagonals D1 bit r+c and D2 bit c-r+n-1), weight 1; the state is
 * propagated first. */
int orc_fullsweep_subtree(unsigned n, uint32_t P, uint32_t C, uint64_t D1, uint64_t D2,
                          uint64_t* sols, uint64_t* nodes);
/* Propagates one state and writes its branch children (P, C, D1, D2 as four
 * u64 each, at most cap of them); returns the number of children (0 when the
 * state is dead or complete) and the propagated placements in *prop_nodes. */
int orc_fullsweep_children(unsigned n, uint32_t P, uint32_t C, uint64_t D1, uint64_t D2,
                           uint64_t* out, unsigned cap, uint64_t* prop_nodes);

/* Serialize a pool in the reference's .nqsf layout (nq_utils.cu:145-193)
 * into buf (capacity cap bytes). Returns bytes written or 0 if cap is too
 * small. flags is the 64-bit compile-flag word the header carries. */
uint64_t orc_nqsf_write(unsigned n, unsigned lock, const uint8_t* pool,
                        uint64_t count, uint64_t flags, uint8_t* buf, uint64_t cap);

#ifdef __cplusplus
}
#endif
#endif
