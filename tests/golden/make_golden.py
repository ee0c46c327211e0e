"""Writes tests/golden/counts.json: oracle counts for seeded graphs that are
too slow to recompute inside a GPU test run.  Run from the repo root:
    python tests/golden/make_golden.py
The oracle itself is pinned by tests/test_oracle.py and
tests/test_oracle_pins.py (brute force, closed forms, per-root attribution);
these fixtures freeze its output for the larger graphs.
    python tests/golden/make_golden.py --rmat22
writes tests/golden/rmat22.json (the benchmark graph: CSR hashes and root
windows of its first quarter and TC/P13 over the whole quarter, about ten
minutes on 8 cores).
    python tests/golden/make_golden.py --chunglu
writes tests/golden/chunglu.json (the induced-motif workload graph).
    python tests/golden/make_golden.py --batch2048
writes tests/golden/batch2048.json (a sample of the batched workload)."""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

import jupiter_amd as J  # noqa: E402
import oracle_lib as O  # noqa: E402

CLIQUE = [0, 13, 14, 15, 16]


# Root windows of the R-MAT scale-22 graph (edge factor 16, seed 1) inside the
# benchmark's quarter [0, 2^20), chosen so that together they hold roots of
# every clique tier (DAG out-degree <= 64, 128, 256, 512, 1024) and the oracle
# finishes each in under a minute.
RMAT22_WINDOWS = [(917500, 918600), (600000, 604096), (300, 302), (1015000, 1016000), (1032000, 1033300),
                  (4000, 4010)]


def csr_sha256(rowptr, col):
    import hashlib
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(rowptr, dtype=np.int64).tobytes())
    h.update(np.ascontiguousarray(col, dtype=np.int32).tobytes())
    return h.hexdigest()


def rmat22():
    """tests/golden/rmat22.json: hashes of the host-built symmetric and DAG
    CSR, per-tier root counts and oracle TC/P13/P14 counts of the windows."""
    g = J.rmat(22, 16, seed=1)
    d = g.orient()
    dd = np.diff(d.rowptr)
    ent = {"scale": 22, "edge_factor": 16, "abcd": [0.57, 0.19, 0.19, 0.05], "seed": 1, "nv": g.nv,
           "sym_nnz": g.nnz, "dag_nnz": d.nnz, "dag_max_deg": int(dd.max()),
           "sym_sha256": csr_sha256(g.rowptr, g.col), "dag_sha256": csr_sha256(d.rowptr, d.col), "windows": []}
    for a, b in RMAT22_WINDOWS:
        x = dd[a:b]
        w = {"roots": [a, b], "tier_roots": {str(hi): int(((x > lo) & (x <= hi)).sum())
                                              for lo, hi in ((0, 64), (64, 128), (128, 256), (256, 512), (512, 1024))}}
        for p in (0, 13, 14):
            t = time.time()
            w[J.PATTERN_NAMES[p]] = O.count(d.rowptr, d.col, p, a, b)
            print("rmat22", (a, b), J.PATTERN_NAMES[p], w[J.PATTERN_NAMES[p]], f"{time.time() - t:.1f}s", flush=True)
        ent["windows"].append(w)
    # the whole first quarter for TC and P13 (about 7 minutes on 6 cores);
    # P14 over the quarter would take hours here
    q = g.nv // 4
    ent["quarter0"] = {}
    for p in (0, 13):
        t = time.time()
        ent["quarter0"][J.PATTERN_NAMES[p]] = O.count(d.rowptr, d.col, p, 0, q)
        print("rmat22 quarter 0", J.PATTERN_NAMES[p], ent["quarter0"][J.PATTERN_NAMES[p]],
              f"{time.time() - t:.1f}s", flush=True)
    with open(os.path.join(HERE, "rmat22.json"), "w") as f:
        json.dump(ent, f, indent=1, sort_keys=True)


def chung_lu_config():
    """tests/golden/chunglu.json: the induced-motif workload graph (Chung-Lu,
    3.8M vertices, 16.5M target edges, gamma 2.3, max weight 800, seed 1):
    its size, a CSR hash and the oracle's non-induced counts the induced
    motifs are made of (P1, P2, P13, 3-star, tailed triangle)."""
    g = J.chung_lu(3_800_000, 16_500_000, 2.3, 800.0, seed=1)
    d = g.orient()
    ent = {"n": 3_800_000, "target_edges": 16_500_000, "gamma": 2.3, "max_weight": 800.0, "seed": 1,
           "nv": g.nv, "sym_nnz": g.nnz, "max_deg": int(g.degrees().max()), "sym_sha256": csr_sha256(g.rowptr, g.col),
           "counts": {}}
    c = ent["counts"]
    for name, fn in (("P1", lambda: O.count(g.rowptr, g.col, 1)), ("P2", lambda: O.count(g.rowptr, g.col, 2)),
                     ("P13", lambda: O.count(d.rowptr, d.col, 13)), ("TC", lambda: O.count(d.rowptr, d.col, 0)),
                     ("STAR3", lambda: O.vertex_sums(g.rowptr, g.col, 0)),
                     ("TAILTRI", lambda: O.vertex_sums(g.rowptr, g.col, 1))):
        t = time.time()
        c[name] = fn()
        print("chunglu", name, c[name], f"{time.time() - t:.1f}s", flush=True)
    # the vertex-induced motifs by the exact inverse conversion
    S, T, C, D, K = c["STAR3"], c["TAILTRI"], c["P1"], c["P2"], c["P13"]
    c["I_K4"] = K
    c["I_DIAMOND"] = D - 6 * K
    c["I_C4"] = C - D + 3 * K
    c["I_TAILTRI"] = T - 4 * D + 12 * K
    c["I_STAR3"] = S - T + 2 * D - 4 * K
    with open(os.path.join(HERE, "chunglu.json"), "w") as f:
        json.dump(ent, f, indent=1, sort_keys=True)


def config4_sample():
    """tests/golden/batch2048.json: oracle counts for a sample of the batched
    workload's 2,048 graphs (the 16 smallest and 4 of median size), each with
    its own pattern, so a GPU run of the whole batch can be checked against
    the oracle without recomputing it there."""
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    import bench
    specs = bench.config4_specs()
    order = sorted(range(len(specs)), key=lambda i: specs[i][1])
    pick = sorted(order[:16] + order[len(order) // 2: len(order) // 2 + 4])
    out = {"n_graphs": len(specs), "sample": []}
    for i in pick:
        g = bench.config4_graph(specs[i])
        p = specs[i][4]
        t = time.time()
        if p in ("TC", "P13"):
            d = g.orient()
            c = O.count(d.rowptr, d.col, J.PATTERNS[p])
        else:
            c = O.count(g.rowptr, g.col, J.PATTERNS[p])
        out["sample"].append({"index": i, "kind": specs[i][0], "n": specs[i][1], "pattern": p, "sym_nnz": g.nnz,
                              "count": c})
        print("batch2048", i, specs[i][:3], p, c, f"{time.time() - t:.1f}s", flush=True)
    with open(os.path.join(HERE, "batch2048.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def main():
    if "--batch2048" in sys.argv:
        config4_sample()
        return
    if "--rmat22" in sys.argv:
        rmat22()
        return
    if "--chunglu" in sys.argv:
        chung_lu_config()
        return
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
