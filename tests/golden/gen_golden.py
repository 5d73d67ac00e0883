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

Run from the repo root:  python tests/golden/gen_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402

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


def main():
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
    print("golden fixtures written")


if __name__ == "__main__":
    main()
