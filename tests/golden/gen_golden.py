"""Writes the golden fixtures under tests/golden/.

Two kinds of data live here:
  * reference data, typed in from the reference's own files:
      known_totals.json  -- the known-answer table, test.cu:18 (N=1..27)
      pool_sizes.json    -- (N, lock, count) from the pool file names,
                            test.cu:88-104 (full pools, all row-0 columns)
      node_totals.json   -- full-tree placement counts N=8..18 computed by an
                            independent textbook DFS (not reference data; the
                            reference never counts nodes)
  * oracle-written vectors (regression fixtures for the GPU path):
      n12_d3_prefixes.npz -- all 756 depth-3 prefixes of N=12 with each
                             prefix's solution and node count
      n16_d5_sample.npz   -- 512 seeded depth-5 prefixes of N=16 with counts
      c3_batch.json       -- the 512-solve mixed batch (seeded, splitmix64):
                             each solve's n, prefix, count and nodes below it
      fullsweep.json      -- full DoubleSweep search (oracle/nq_fullsweep.c):
                             solutions and placements for N=1..18 (N >= 15 take
                             the oracle seconds to minutes on 8 threads)

Run from the repo root:  python tests/golden/gen_golden.py [fullsweep]
(with `fullsweep`, only fullsweep.json is written).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "gpu-n-queens_amd"))
import oracle  # noqa: E402
from nqm import workloads  # noqa: E402

KNOWN = [1, 0, 0, 2, 10, 4, 40, 92, 352, 724, 2680, 14200, 73712, 365596, 2279184, 14772512,
         95815104, 666090624, 4968057848, 39029188884, 314666222712, 2691008701644,
         24233937684440, 227514171973736, 2207893435808352, 22317699616364044,
         234907967154122528]

POOLS = [(10, 7, 6358), (11, 7, 30040), (12, 8, 152604), (13, 9, 802192), (14, 10, 4568428),
         (15, 11, 27851188), (16, 9, 44903824), (17, 8, 35255116), (18, 8, 75913592),
         (19, 7, 26512942), (20, 7, 45562852), (21, 7, 75580634), (22, 6, 13145292),
         (23, 6, 18908302), (24, 6, 26670584), (25, 6, 36961170), (26, 6, 50409604)]

NODES = {8: 2056, 9: 8393, 10: 35538, 11: 166925, 12: 856188, 13: 4674889, 14: 27358552,
         15: 171129071, 16: 1141190302, 17: 8017021931, 18: 59365844490}


def extensions(n, prefix, depth):
    """Valid extensions of `prefix` to `depth` rows, and the number of
    placements made on the way (rows len(prefix)..depth-1)."""
    mask = (1 << n) - 1
    cols = ld = rd = 0
    for c in prefix:
        b = 1 << c
        cols |= b
        ld = ((ld | b) << 1) & mask
        rd = (rd | b) >> 1
    out, placed = [], [0]

    def rec(row, cols, ld, rd, pre):
        if row == depth:
            out.append(pre)
            return
        f = ~(cols | ld | rd) & mask
        while f:
            b = f & -f
            f ^= b
            placed[0] += 1
            rec(row + 1, cols | b, ((ld | b) << 1) & mask, (rd | b) >> 1, pre + [b.bit_length() - 1])

    rec(len(prefix), cols, ld, rd, list(prefix))
    return out, placed[0]


def oracle_solve(n, prefix):
    """(solutions, nodes below prefix) with the threaded oracle pool counter."""
    k = len(prefix)
    if k == n:
        return 1, 0
    d = min(n - 1, k + 3)
    if d <= k:
        return oracle.count_prefix(n, list(prefix))
    pool, placed = extensions(n, prefix, d)
    if not pool:
        return 0, placed
    tot, nd = oracle.count_pool(n, d, np.array(pool, np.uint8))
    return tot, nd + placed


def c3_fixture():
    solves = workloads.mixed_batch()
    cache = {}
    rows = []
    for n, pre in solves:
        key = (n, pre)
        if key not in cache:
            cache[key] = oracle_solve(n, pre)
        c, nd = cache[key]
        rows.append({"n": n, "prefix": list(pre), "count": c, "nodes": nd})
    return {"seed": "0x5EED0512", "generator": "nqm.workloads.mixed_batch (splitmix64)",
            "solves": rows}


def fullsweep_fixture(max_n=18, threads=8):
    rows = {}
    for n in range(1, max_n + 1):
        s, d = oracle.fullsweep_count(n, threads)
        assert s == KNOWN[n - 1], (n, s)
        rows[str(n)] = {"solutions": s, "nodes": d}
        print("fullsweep", n, s, d, flush=True)
    with open(os.path.join(HERE, "fullsweep.json"), "w") as f:
        json.dump({"source": "oracle/nq_fullsweep.c (mirror-halved search; nodes = placements executed)",
                   "counts": rows}, f, indent=1)


def main():
    if sys.argv[1:] == ["fullsweep"]:
        fullsweep_fixture()
        return
    with open(os.path.join(HERE, "known_totals.json"), "w") as f:
        json.dump({"source": "test.cu:18", "totals": {str(i + 1): v for i, v in enumerate(KNOWN)}},
                  f, indent=1)
    with open(os.path.join(HERE, "pool_sizes.json"), "w") as f:
        json.dump({"source": "test.cu:88-104 (file names n<N>_<count>_lk<lock>_valid.nqsb)",
                   "mirror": False,
                   "pools": [{"n": n, "lock": l, "count": c} for n, l, c in POOLS]}, f, indent=1)
    with open(os.path.join(HERE, "node_totals.json"), "w") as f:
        json.dump({"source": "independent textbook DFS (placements incl. row 0)",
                   "nodes": {str(k): v for k, v in NODES.items()}}, f, indent=1)

    pool = oracle.pool_generate(12, 3, mirror=False)
    tot, nd, ps, pn = oracle.count_pool(12, 3, pool, per_prefix=True)
    assert tot == 14200
    np.savez_compressed(os.path.join(HERE, "n12_d3_prefixes.npz"), pool=pool, sols=ps, nodes=pn)

    rng = np.random.default_rng(20261203)
    full = oracle.pool_generate(16, 5, mirror=False)
    idx = np.sort(rng.choice(full.shape[0], 512, replace=False))
    sample = full[idx]
    _, _, ps, pn = oracle.count_pool(16, 5, sample, per_prefix=True)
    np.savez_compressed(os.path.join(HERE, "n16_d5_sample.npz"), pool=sample, sols=ps, nodes=pn,
                        index=idx)
    with open(os.path.join(HERE, "c3_batch.json"), "w") as f:
        json.dump(c3_fixture(), f, indent=0)
    fullsweep_fixture()
    print("golden fixtures written")


if __name__ == "__main__":
    main()
