"""The induced-motif, batched and serving workloads at their own sizes.

* Induced 4-vertex motifs on the Chung-Lu graph of 3.8M vertices (16.5M
  target edges, gamma 2.3, max weight 800): the host build equals the frozen
  one (SHA-256), and every non-induced count the motifs are made of, and the
  motifs themselves, equal the oracle's frozen counts.
* The batch of 2,048 graphs (half Erdos-Renyi, half Barabasi-Albert, n
  log-uniform in [1k, 262k], one pattern each from {TC, P13, P1, P2}): the
  captured and replayed batch equals one plain launch per query on every
  graph, and a sample equals the oracle's frozen counts.
* Serving: one R-MAT graph per scale 18..24; every pattern the serving
  benchmark draws at that scale, submitted together to a server, gives the
  same count as alone.
"""
import hashlib
import json
import os
import time

import numpy as np
import pytest

import bench
import jupiter_amd as J

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(msg):
    print(f"[case] {msg}", flush=True)


def test_induced_motifs_full_size(ctx):
    cl = json.load(open(os.path.join(GOLD, "chunglu.json")))
    g = J.chung_lu(cl["n"], cl["target_edges"], cl["gamma"], cl["max_weight"], seed=cl["seed"])
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(g.rowptr, dtype=np.int64).tobytes())
    h.update(np.ascontiguousarray(g.col, dtype=np.int32).tobytes())
    assert g.nv == cl["nv"] and g.nnz == cl["sym_nnz"] and h.hexdigest() == cl["sym_sha256"]
    G = ctx.upload(g)
    del g
    try:
        for p in ("P1", "P2", "P13", "TC", "STAR3", "TAILTRI", "I_K4", "I_DIAMOND", "I_C4", "I_TAILTRI", "I_STAR3"):
            _say(f"chung-lu 3.8M {p}")
            assert G.count(p)[0] == cl["counts"][p], p
    finally:
        G.free()


def test_batch_2048_full_size(ctx):
    gold = json.load(open(os.path.join(GOLD, "batch2048.json")))
    specs = bench.config4_specs()
    assert len(specs) == gold["n_graphs"]
    Gs = []
    t0 = time.time()
    try:
        for i, spec in enumerate(specs):
            Gs.append(ctx.upload(bench.config4_graph(spec)))
            if (i + 1) % 256 == 0:
                _say(f"batch: {i + 1} graphs built and uploaded ({time.time() - t0:.0f}s)")
        queries = [(G, s[4], 0, None) for G, s in zip(Gs, specs)]
        _say("batch: plain launches")
        plain = [G.count(p)[0] for G, p, _, _ in queries]
        for ent in gold["sample"]:
            assert plain[ent["index"]] == ent["count"], ent
        _say("batch: captured batch, two replays")
        with ctx.prepare_batch(queries) as B:
            for _ in range(2):
                got, _ = B.run()
                assert got == plain
        _say("batch: eager batch")
        got, _ = ctx.count_batch(queries)
        assert got == plain
    finally:
        for G in Gs:
            G.free()


@pytest.mark.parametrize("scale", range(18, 25))
def test_serving_scales_equal_alone(ctx, scale):
    pats = [p for p in bench.SERVE_PATTERNS if bench.SERVE_CAP[p] >= scale] + ["TC"]
    _say(f"serving scale {scale}: R-MAT on the device, patterns {pats}")
    G = ctx.rmat_graph(scale, 16, 0.57, 0.19, 0.19, seed=bench.splitmix(bench.SEED, scale))
    try:
        alone = {}
        for p in pats:
            _say(f"serving scale {scale}: {p} alone")
            alone[p] = G.count(p)[0]
        _say(f"serving scale {scale}: served together")
        with J.Server(ctx) as srv:
            tickets = {p: srv.submit(G, p) for p in pats}
            served = {p: srv.wait(t)[0] for p, t in tickets.items()}
        assert served == alone
    finally:
        G.free()
