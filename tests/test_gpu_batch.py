"""Batched persistent launch: every query's count equals its single-query
count and the oracle (bit-exact), in any mix of graphs and patterns."""
import numpy as np
import pytest

import jupiter_amd as J
import oracle_lib as O
from gen import hub_graph

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)


def oracle_count(csr, dag, p):
    pid = J.PATTERNS[p]
    if pid in (0, 13, 14, 15, 16):
        return O.count(dag.rowptr, dag.col, pid)
    return O.count(csr.rowptr, csr.col, pid)


def test_mixed_batch_matches_oracle_and_alone(ctx):
    rng = np.random.default_rng(5)
    graphs, queries, expect = [], [], []
    pats = ["TC", "P13", "P1", "P2"]
    for i in range(24):
        n = int(np.exp(rng.uniform(np.log(300), np.log(3000))))
        if i % 2 == 0:
            g = J.erdos_renyi(n, float(rng.uniform(8, 24)), seed=100 + i)
        else:
            g = J.barabasi_albert(n, int(rng.integers(3, 8)), seed=100 + i)
        d = g.orient()
        G = ctx.upload(g)
        graphs.append(G)
        p = pats[int(rng.integers(0, 4))]
        queries.append((G, p, 0, None))
        expect.append(oracle_count(g, d, p))
    # a few more pattern kinds, a root range, and wide clique roots
    g = J.rmat(11, 16, seed=3)
    d = g.orient()
    G = ctx.upload(g)
    graphs.append(G)
    for p in ["P3", "P5", "P7", "P9", "P12", "P14", "P15", "I_DIAMOND", "STAR3"]:
        queries.append((G, p, 0, None))
        if p in ("I_DIAMOND", "STAR3"):
            expect.append(G.count(p)[0])
        else:
            expect.append(oracle_count(g, d, p))
    queries.append((G, "P13", 100, 900))
    expect.append(O.count(d.rowptr, d.col, 13, 100, 900))
    h = hub_graph(700, 0.03, seed=7)
    hd = h.orient()
    H = ctx.upload(h)
    graphs.append(H)
    for p in ["TC", "P13", "P14"]:
        queries.append((H, p, 0, None))
        expect.append(oracle_count(h, hd, p))
    _say(f"batch of {len(queries)} queries")
    got, ms = ctx.count_batch(queries)
    assert got == expect
    # same results alone, and in another order
    alone = [Gq.count(p, None if rb == 0 and re is None else (rb, re))[0] for Gq, p, rb, re in queries]
    assert alone == expect
    order = list(range(len(queries)))[::-1]
    got2, _ = ctx.count_batch([queries[i] for i in order])
    assert [got2[order.index(i)] for i in range(len(queries))] == expect
    for G in graphs:
        G.free()


def test_batch_edge_cases(ctx):
    assert ctx.count_batch([]) == ([], 0.0)
    g = J.rmat(8, 8, seed=1)
    G = ctx.upload(g)
    with pytest.raises(J.JgpmError) as e:
        ctx.count_batch([(G, 99, 0, None)])
    assert e.value.status == -2
    with pytest.raises(J.JgpmError):
        ctx.count_batch([(G, "TC", 5, 2)])
    got, _ = ctx.count_batch([(G, "TC", 3, 3), (G, "P2", 0, 0)])
    assert got == [0, 0]
    G.free()


def test_batch_builds_edge_triangles_after_another_build(ctx):
    """A batch whose plan builds a graph's per-edge triangle array right after
    a single query built another graph's (the builder's queue counter left at
    that graph's edge count) counts every edge: P3, P7 and P11 in batches on
    fresh graphs equal the oracle."""
    a = J.rmat(10, 16, seed=41)
    b = J.barabasi_albert(1500, 6, seed=42)
    Ga, Gb = ctx.upload(a), ctx.upload(b)
    try:
        _say("batch after a single-query triangle build")
        assert Ga.count("P3")[0] == O.count(a.rowptr, a.col, J.PATTERNS["P3"])
        got, _ = ctx.count_batch([(Gb, "P3", 0, None), (Gb, "P7", 0, None)])
        assert got == [O.count(b.rowptr, b.col, J.PATTERNS["P3"]), O.count(b.rowptr, b.col, J.PATTERNS["P7"])]
        Gc = ctx.upload(a)
        try:
            with ctx.prepare_batch([(Gc, "P11", 0, None)]) as B:
                got, _ = B.run()
            assert got == [O.count(a.rowptr, a.col, J.PATTERNS["P11"])]
        finally:
            Gc.free()
    finally:
        Ga.free()
        Gb.free()
