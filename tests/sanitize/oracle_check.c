/* The CPU oracle (oracle/nq_oracle.c, nq_symmetry.c) under ASan/UBSan on small
 * boards, against known answers: totals of test.cu:18, the full-tree node
 * numbers, the pool sizes of test.cu:88-104 and OEIS A002562. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../oracle/nq_oracle.h"

static int bad(const char* what, unsigned n) {
    fprintf(stderr, "oracle_check: %s (n=%u)\n", what, n);
    return 1;
}

int main(void) {
    /* test.cu:18, N = 1..12 */
    static const uint64_t total[13] = {0, 1, 0, 0, 2, 10, 4, 40, 92, 352, 724, 2680, 14200};
    /* full-tree placements, N = 8..12 (tests/golden/node_totals.json) */
    static const uint64_t nodes[13] = {0, 0, 0, 0, 0, 0, 0, 0, 2056, 8393, 35538, 166925, 856188};
    /* OEIS A002562, N = 1..12 */
    static const uint64_t unique[13] = {0, 1, 0, 0, 1, 2, 1, 6, 12, 46, 92, 341, 1787};
    for (unsigned n = 1; n <= 12; ++n) {
        uint64_t lv[33] = {0};
        orc_level_counts(n, n, lv);
        uint64_t sum = 0;
        for (unsigned k = 1; k <= n; ++k) sum += lv[k];
        if ((n >= 8 && sum != nodes[n]) || lv[n] != total[n]) return bad("level counts", n);
        for (unsigned depth = 1; depth < n && depth <= 4; ++depth) {
            for (int mirror = 0; mirror <= 1; ++mirror) {
                const uint64_t cnt = orc_pool_count(n, depth, mirror, NULL);
                uint8_t* pool = malloc(cnt * depth + 1);
                if (!pool || orc_pool_generate(n, depth, mirror, pool, cnt) != cnt) return bad("pool", n);
                uint64_t t = 0, nd = 0;
                if (cnt && orc_count_pool(n, depth, pool, cnt, mirror, 3, &t, &nd, NULL, NULL) != 0)
                    return bad("count", n);
                if (cnt && t != total[n]) return bad("pool total", n);
                uint8_t buf[1 << 16];
                if (cnt * depth < sizeof buf - 64 && orc_nqsf_write(n, depth, pool, cnt, 3, buf, sizeof buf) == 0)
                    return bad("nqsf write", n);
                free(pool);
            }
        }
        uint64_t u = 0, tot = 0, sn = 0;
        if (orc_sym_count(n, &u, &tot, &sn) != 0 || u != unique[n] || tot != total[n]) return bad("symmetric", n);
        if (n <= 10) {
            uint64_t bu = 0, bt = 0;
            if (orc_sym_brute(n, &bu, &bt) != 0 || bu != unique[n] || bt != total[n]) return bad("brute", n);
        }
    }
    /* test.cu:88-104: (10, lock 7) holds 6358 prefixes (all row-0 columns) */
    if (orc_pool_count(10, 7, 0, NULL) != 6358) return bad("pool size 10/7", 10);
    unsigned lock = 0;
    uint64_t cnt = 0;
    if (orc_lock_depth(12, 20000, 1, &lock, &cnt) != 0 || cnt != orc_pool_count(12, lock, 1, NULL))
        return bad("lock depth", 12);
    printf("oracle_check: ok (N=1..12)\n");
    return 0;
}
