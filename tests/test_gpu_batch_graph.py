"""Prepared batches (all launches of a batched count captured once in a HIP
graph) and the batch autotuner.

Bar: bit-exact.  Every replay, every tuning candidate and every knob setting
gives the counts of the eager batch and of each query alone; replays are
interleaved with eager work and synchronisation, and each is checked."""
import numpy as np
import pytest

import jupiter_amd as J
import oracle_lib as O
from gen import hub_graph

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _mix(ctx):
    """Small ER/BA graphs with TC/P13/P1/P2, an R-MAT with more P-ids and a
    root range, a hub graph (block-per-hub launches up to the 512/1024-wide
    tiers) and K_1100 (global-row tier)."""
    rng = np.random.default_rng(11)
    graphs, queries = [], []
    pats = ["TC", "P13", "P1", "P2"]
    for i in range(10):
        n = int(np.exp(rng.uniform(np.log(300), np.log(2500))))
        g = J.erdos_renyi(n, float(rng.uniform(8, 24)), seed=200 + i) if i % 2 == 0 else \
            J.barabasi_albert(n, int(rng.integers(3, 8)), seed=200 + i)
        G = ctx.upload(g)
        graphs.append(G)
        queries.append((G, pats[i % 4], 0, None))
    g = J.rmat(10, 16, seed=9)
    G = ctx.upload(g)
    graphs.append(G)
    for p in ["P3", "P5", "P8", "P11", "P14", "P15"]:
        queries.append((G, p, 0, None))
    queries.append((G, "P13", 50, 600))
    h = hub_graph(900, 0.03, seed=8)
    H = ctx.upload(h)
    graphs.append(H)
    for p in ["TC", "P13", "P14"]:
        queries.append((H, p, 0, None))
    src, dst = np.triu_indices(1100, 1)
    K = ctx.upload(J.CSR.from_edges(1100, src, dst))
    graphs.append(K)
    queries.append((K, "TC", 0, None))
    queries.append((K, "P13", 0, 3))
    return graphs, queries


def _alone(queries):
    return [G.count(p, None if rb == 0 and re is None else (rb, re))[0] for G, p, rb, re in queries]


def _restore(ctx):
    ctx.set_tuning(**{f: -1 for f in ctx.tuning()})


def test_prepared_batch_replays(ctx):
    graphs, queries = _mix(ctx)
    _say(f"prepared batch of {len(queries)} queries")
    eager, _ = ctx.count_batch(queries)
    assert eager == _alone(queries)
    from math import comb
    assert eager[-2] == comb(1100, 3)
    B = ctx.prepare_batch(queries)
    try:
        for r in range(3):
            got, ms = B.run()
            assert got == eager, f"replay {r}"
            assert ms > 0
            # eager work and synchronisation between replays
            assert ctx.count_batch(queries[:4])[0] == eager[:4]
            assert graphs[0].count("TC")[0] == eager[0]
        # a graph a live batch launches on cannot be freed
        with pytest.raises(J.JgpmError) as e:
            graphs[0].free()
        assert e.value.status == -8
    finally:
        B.free()
    for G in graphs:
        G.free()


@pytest.mark.parametrize("hub,rchunk,echunk,bpc", [(64, 1, 1, 1), (64, 64, 32, 0), (128, 4, 2, 1)])
def test_batch_knobs_keep_counts(ctx, hub, rchunk, echunk, bpc):
    graphs, queries = _mix(ctx)
    ref = _alone(queries)
    try:
        ctx.set_tuning(batch_hub_deg=hub, batch_root_chunk=rchunk, batch_edge_chunk=echunk,
                       batch_blocks_per_cu=bpc if bpc else -1)
        _say(f"knobs hub={hub} rchunk={rchunk} echunk={echunk} bpc={bpc}")
        assert ctx.count_batch(queries)[0] == ref
        with ctx.prepare_batch(queries) as B:
            assert B.run()[0] == ref
    finally:
        _restore(ctx)
    for G in graphs:
        G.free()


def test_autotune_batch(ctx):
    graphs, queries = _mix(ctx)
    ref = _alone(queries)
    try:
        best, ms = ctx.autotune_batch(queries, reps=1)
        _say(f"autotuned {best} {ms:.2f} ms")
        assert best["batch_hub_deg"] in (64, 128)
        assert best["batch_root_chunk"] in (4, 16, 64)
        assert best["batch_edge_chunk"] in (2, 8, 32)
        assert ms > 0
        assert ctx.tuning() == best  # installed
        assert ctx.count_batch(queries)[0] == ref
        with ctx.prepare_batch(queries) as B:
            assert B.run()[0] == ref
    finally:
        _restore(ctx)
    for G in graphs:
        G.free()


def test_prepared_batch_rejects(ctx):
    g = J.rmat(8, 8, seed=2)
    G = ctx.upload(g)
    with pytest.raises(J.JgpmError) as e:
        ctx.prepare_batch([(G, "I_DIAMOND", 0, None)])  # composite ids are not run by the kernel
    assert e.value.status == -2
    with pytest.raises(J.JgpmError):
        ctx.prepare_batch([])
    G.free()  # no batch holds it


def test_prepared_batch_keeps_its_wedge_counters(ctx):
    """4-cycle queries of a prepared batch use counters owned by the batch:
    a later plain 4-cycle query on a larger graph (which grows the context's
    counters) does not move them, and replays after it keep their counts."""
    small = [ctx.upload(J.erdos_renyi(2000, 12, seed=300 + i)) for i in range(3)]
    queries = [(G, "P1", 0, None) for G in small] + [(small[0], "P1", 100, 1500)]
    want = _alone(queries)
    with ctx.prepare_batch(queries) as B:
        _say("prepared P1 batch, first replay")
        assert B.run()[0] == want
        big = J.erdos_renyi(200_000, 16, seed=310)
        G = ctx.upload(big)
        _say("plain P1 on 200k vertices")
        assert G.count("P1")[0] == O.count(big.rowptr, big.col, J.PATTERNS["P1"])
        for rep in range(2):
            _say(f"replay {rep} after the context's counters grew")
            assert B.run()[0] == want
        G.free()
    for G in small:
        G.free()
