"""GPU parity of the clique engine (TC, P13..P16) against the CPU oracle.

Bar: bit-exact uint64 counts (integer work, tolerance 0)."""
import json
import os

import numpy as np
import pytest

import jupiter_amd as J
import oracle_lib as O
from gen import hub_graph

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "counts.json")))
CLIQUES = ["TC", "P13", "P14", "P15", "P16"]


def _say(msg):
    print(f"[case] {msg}", flush=True)


@pytest.mark.parametrize("scale,seed", [(8, 1), (10, 3), (11, 5)])
def test_rmat_small_vs_oracle(ctx, scale, seed):
    g = J.rmat(scale, 16, seed=seed)
    d = g.orient()
    G = ctx.upload(g)
    for p in CLIQUES:
        _say(f"rmat{scale} seed{seed} {p}")
        got, _ = G.count(p)
        assert got == O.count(d.rowptr, d.col, J.PATTERNS[p]), p
    G.free()


@pytest.mark.parametrize("key", ["rmat12_s7", "rmat14_s1"])
def test_rmat_golden(ctx, key):
    ent = GOLD[key]
    g = J.rmat(ent["scale"], 16, seed=ent["seed"])
    assert g.nnz == ent["sym_nnz"]
    G = ctx.upload(g)
    q = g.nv // 4
    for p in CLIQUES:
        _say(f"{key} {p}")
        got, ms = G.count(p)
        assert got == ent["counts"][p], p
        quarters = [G.count(p, (i * q, (i + 1) * q))[0] for i in range(4)]
        assert quarters == ent["quarters"][p], p
        assert sum(quarters) == got
        assert G.work_bytes(p) == ent["work_bytes"][p], p
    G.free()


@pytest.mark.parametrize("core,p", [(100, 0.3), (200, 0.15), (400, 0.06), (700, 0.03), (1500, 0.012)])
def test_wide_root_tiers(ctx, core, p):
    """Roots with DAG out-degree 100..1500 run in the LDS and global-row tiers."""
    g = hub_graph(core, p, seed=core)
    d = g.orient()
    assert d.degrees()[0] == core
    G = ctx.upload(g)
    for pat in CLIQUES:
        _say(f"hub core={core} {pat}")
        got, _ = G.count(pat)
        assert got == O.count(d.rowptr, d.col, J.PATTERNS[pat]), pat
        assert G.count(pat, (0, 1))[0] == O.count(d.rowptr, d.col, J.PATTERNS[pat], 0, 1)
    G.free()


@pytest.mark.parametrize("n", [1, 2, 3, 4, 7, 8, 12, 40, 70, 150, 300, 600, 1100])
def test_complete_graph_closed_form(ctx, n):
    """K_n: vertex 0 has DAG out-degree n-1 and every row i holds n-2-i
    members, so n = 150 / 300 / 600 run the block-cooperative compressed
    rows (|S| > 64) and, for |S| > 512, the word-recompute fallback."""
    from math import comb
    src, dst = np.triu_indices(n, 1)
    g = J.CSR.from_edges(n, src, dst)
    G = ctx.upload(g)
    pats = [("TC", 3), ("P13", 4), ("P14", 5), ("P15", 6), ("P16", 7)]
    if n >= 300:
        pats = pats[:3]
    elif n >= 150:
        pats = pats[:4]
    for pat, k in pats:
        _say(f"K{n} {pat}")
        assert G.count(pat)[0] == comb(n, k), (n, pat)
    G.free()


def test_edge_cases(ctx):
    # no edges, a single edge, empty root range
    g = J.CSR.from_edges(5, [], [])
    G = ctx.upload(g)
    for pat in CLIQUES:
        assert G.count(pat)[0] == 0
    G.free()
    g = J.CSR.from_edges(2, [0], [1])
    G = ctx.upload(g)
    assert G.count("TC")[0] == 0
    assert G.count("TC", (1, 1))[0] == 0
    G.free()


def test_bad_arguments(ctx):
    g = J.rmat(6, 4, seed=2)
    G = ctx.upload(g)
    with pytest.raises(J.JgpmError) as e:
        G.count(99)
    assert e.value.status == -2
    with pytest.raises(J.JgpmError) as e:
        G.count("TC", (5, 2))
    assert e.value.status == -1
    with pytest.raises(J.JgpmError):
        G.count("TC", (0, g.nv + 1))
    G.free()


def test_gallop_ratio_does_not_change_counts(ctx):
    g = J.rmat(12, 16, seed=7)
    G = ctx.upload(g)
    ent = GOLD["rmat12_s7"]
    for ratio in (1, 4, 1000000):
        ctx.set_tuning(gallop_ratio=ratio)
        for p in ("TC", "P13", "P14"):
            _say(f"ratio {ratio} {p}")
            assert G.count(p)[0] == ent["counts"][p]
    ctx.set_tuning(gallop_ratio=32)
    G.free()


def test_dag_only_upload_uses_topological_order(ctx):
    """A DAG uploaded as such gets its member order from Kahn's algorithm;
    counts match the oracle on roots wide enough for multi-word rows."""
    g = J.rmat(13, 16, seed=9)
    d = g.orient()
    assert d.degrees().max() > 64
    G = ctx.upload(d, flags=J.DAG)
    for p in ["TC", "P13", "P14"]:
        _say(f"dag-only {p}")
        assert G.count(p)[0] == O.count(d.rowptr, d.col, J.PATTERNS[p]), p
    G.free()


def test_roots_wider_than_2048(ctx):
    """Roots with DAG out-degree above 2048 (beyond the widest row tier) run
    the level nest in global scratch: a 2,600-wide sparse hub, and K_2100
    (52 such roots) for triangles; counts against the oracle / closed form,
    whole graph, the hub's root alone and through the serving loop."""
    from math import comb
    g = hub_graph(2600, 0.006, seed=26)
    d = g.orient()
    assert d.degrees()[0] == 2600
    G = ctx.upload(g)
    try:
        for pat in CLIQUES:
            _say(f"hub core=2600 {pat}")
            want = O.count(d.rowptr, d.col, J.PATTERNS[pat])
            assert G.count(pat)[0] == want, pat
            assert G.count(pat, (0, 1))[0] == O.count(d.rowptr, d.col, J.PATTERNS[pat], 0, 1), pat
        with J.Server(ctx) as srv:
            t = srv.submit(G, "P14")
            assert srv.wait(t)[0] == O.count(d.rowptr, d.col, J.PATTERNS["P14"])
    finally:
        G.free()
    n = 2100
    src, dst = np.triu_indices(n, 1)
    G = ctx.upload(J.CSR.from_edges(n, src, dst))
    try:
        _say("K2100 TC")
        assert G.count("TC")[0] == comb(n, 3)
    finally:
        G.free()
