"""Writes tests/golden/counts.json: oracle counts for seeded graphs that are
too slow to recompute inside a GPU test run.  Run from the repo root:
    python tests/golden/make_golden.py
The oracle itself is pinned by tests/test_oracle.py (brute force, closed
forms); these fixtures freeze its output for the larger graphs."""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import jupiter_amd as J  # noqa: E402
import oracle_lib as O  # noqa: E402

CLIQUE = [0, 13, 14, 15, 16]


def main():
    out = {}
    for scale, seed, pats in [(14, 1, CLIQUE + [1, 2]), (12, 7, CLIQUE + list(range(1, 13)))]:
        g = J.rmat(scale, 16, seed=seed)
        d = g.orient()
        key = f"rmat{scale}_s{seed}"
        ent = {"scale": scale, "edge_factor": 16, "abcd": [0.57, 0.19, 0.19, 0.05], "seed": seed,
               "nv": g.nv, "sym_nnz": g.nnz, "dag_nnz": d.nnz, "counts": {}, "quarters": {}, "work_bytes": {}}
        q = g.nv // 4
        for p in pats:
            t = time.time()
            if p in CLIQUE:
                ent["counts"][J.PATTERN_NAMES[p]] = O.count(d.rowptr, d.col, p)
                ent["quarters"][J.PATTERN_NAMES[p]] = [O.count(d.rowptr, d.col, p, i * q, (i + 1) * q) for i in range(4)]
                ent["work_bytes"][J.PATTERN_NAMES[p]] = O.work_bytes(d.rowptr, d.col, p)
            else:
                ent["counts"][J.PATTERN_NAMES[p]] = O.count(g.rowptr, g.col, p)
                ent["quarters"][J.PATTERN_NAMES[p]] = [O.count(g.rowptr, g.col, p, i * q, (i + 1) * q) for i in range(4)]
            print(key, J.PATTERN_NAMES[p], ent["counts"][J.PATTERN_NAMES[p]], f"{time.time() - t:.1f}s", flush=True)
        out[key] = ent
    with open(os.path.join(HERE, "counts.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
