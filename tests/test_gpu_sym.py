"""GPU parity of the symmetric-graph patterns P1..P12, the vertex sums and
the induced 4-vertex motifs against the CPU oracle / brute force.  Bar:
bit-exact uint64 counts."""
import itertools
import json
import os
import random
from math import comb

import numpy as np
import pytest

import brute
import graphs
import jupiter_amd as J
import oracle_lib as O
from gen import hub_graph

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "counts.json")))
SYM = [f"P{i}" for i in range(1, 13)]


def _say(msg):
    print(f"[case] {msg}", flush=True)


def induced4(adj):
    """Brute-force vertex-induced 4-vertex motif counts."""
    nv = len(adj)
    out = {"I_C4": 0, "I_DIAMOND": 0, "I_TAILTRI": 0, "I_STAR3": 0, "I_K4": 0}
    for quad in itertools.combinations(range(nv), 4):
        es = [(a, b) for a, b in itertools.combinations(quad, 2) if b in adj[a]]
        m = len(es)
        deg = sorted(sum(1 for e in es if v in e) for v in quad)
        if m == 6:
            out["I_K4"] += 1
        elif m == 5:
            out["I_DIAMOND"] += 1
        elif m == 4 and deg == [2, 2, 2, 2]:
            out["I_C4"] += 1
        elif m == 4 and deg == [1, 2, 2, 3]:
            out["I_TAILTRI"] += 1
        elif m == 3 and deg == [1, 1, 1, 3]:
            out["I_STAR3"] += 1
    return out


def test_rmat12_golden(ctx):
    ent = GOLD["rmat12_s7"]
    g = J.rmat(12, 16, seed=7)
    G = ctx.upload(g)
    q = g.nv // 4
    for p in SYM:
        _say(f"rmat12 {p}")
        got, ms = G.count(p)
        assert got == ent["counts"][p], p
        quarters = [G.count(p, (i * q, (i + 1) * q))[0] for i in range(4)]
        assert quarters == ent["quarters"][p], p
    G.free()


@pytest.mark.parametrize("scale,seed", [(7, 2), (9, 4)])
def test_rmat_small_vs_oracle(ctx, scale, seed):
    g = J.rmat(scale, 8, seed=seed)
    G = ctx.upload(g)
    for p in SYM:
        _say(f"rmat{scale} {p}")
        assert G.count(p)[0] == O.count(g.rowptr, g.col, J.PATTERNS[p]), p
        assert G.count(p, (3, g.nv // 2))[0] == O.count(g.rowptr, g.col, J.PATTERNS[p], 3, g.nv // 2), p
    G.free()


def test_p1_wedges_many_vertices(ctx):
    """P1 on graphs with more vertices than the wedge counters allow every
    wave at full occupancy (the wave count is then limited by the counter
    budget): equal to the oracle over the whole range and a window, the
    quarters sum to the whole, and queries repeated and interleaved across
    graphs agree (every launch leaves its counters at zero)."""
    graphs = [J.erdos_renyi(150_000, 64, seed=5), J.barabasi_albert(100_000, 8, seed=6)]
    Gs = [ctx.upload(g) for g in graphs]
    want = []
    for g in graphs:
        _say(f"P1 oracle nv={g.nv}")
        want.append(O.count(g.rowptr, g.col, J.PATTERNS["P1"]))
    for rep in range(2):
        for g, G, w in zip(graphs, Gs, want):
            _say(f"P1 nv={g.nv} rep {rep}")
            assert G.count("P1")[0] == w
    g, G = graphs[0], Gs[0]
    q = g.nv // 4
    assert sum(G.count("P1", (i * q, (i + 1) * q if i < 3 else g.nv))[0] for i in range(4)) == want[0]
    assert G.count("P1", (1000, 90_000))[0] == O.count(g.rowptr, g.col, J.PATTERNS["P1"], 1000, 90_000)
    for G in Gs:
        G.free()


@pytest.mark.parametrize("n", [4, 6, 9, 13])
def test_complete_graph_closed_form(ctx, n):
    src, dst = np.triu_indices(n, 1)
    G = ctx.upload(J.CSR.from_edges(n, src, dst))
    for p in SYM:
        pid = J.PATTERNS[p]
        k, edges = brute.PATTERNS[pid]
        expect = comb(n, k) * int(np.prod(range(1, k + 1))) // brute.automorphisms(k, edges)
        _say(f"K{n} {p}")
        assert G.count(p)[0] == expect, (n, p)
    G.free()


def test_long_lists_use_global_scratch(ctx):
    # sym degrees above the per-wave LDS capacity
    g = hub_graph(300, 0.08, seed=11)
    G = ctx.upload(g)
    for p in ["P1", "P2", "P3", "P4", "P5", "P7", "P8", "P9", "P10", "P11", "P12"]:
        _say(f"hub300 {p}")
        assert G.count(p)[0] == O.count(g.rowptr, g.col, J.PATTERNS[p]), p
    G.free()


@pytest.mark.parametrize("seed", range(3))
def test_induced_motifs_and_vertex_sums_brute_force(ctx, seed):
    rng = random.Random(seed)
    nv = 26
    rp, col, adj = graphs.random_graph(nv, 0.3, rng)
    G = ctx.upload(J.CSR(rp, col))
    want = induced4(adj)
    for name, val in want.items():
        _say(f"induced {name} seed {seed}")
        assert G.count(name)[0] == val, name
    assert G.count("STAR3")[0] == brute.count_pattern(adj, 17)
    assert G.count("TAILTRI")[0] == brute.count_pattern(adj, 18)
    G.free()


def test_induced_identity_on_rmat(ctx):
    ent = GOLD["rmat12_s7"]
    g = J.rmat(12, 16, seed=7)
    G = ctx.upload(g)
    n_s = sum(comb(int(d), 3) for d in g.degrees())
    n_d, n_c, n_k = ent["counts"]["P2"], ent["counts"]["P1"], ent["counts"]["P13"]
    assert G.count("STAR3")[0] == n_s
    n_t = G.count("TAILTRI")[0]
    assert G.count("I_K4")[0] == n_k
    assert G.count("I_DIAMOND")[0] == n_d - 6 * n_k
    assert G.count("I_C4")[0] == n_c - n_d + 3 * n_k
    assert G.count("I_TAILTRI")[0] == n_t - 4 * n_d + 12 * n_k
    assert G.count("I_STAR3")[0] == n_s - n_t + 2 * n_d - 4 * n_k
    G.free()


def test_induced_motifs_chung_lu(ctx):
    """Vertex-induced 4-vertex motifs on power-law Chung-Lu graphs (the family
    of the induced-motif workload: gamma 2.3, max weight 800; the larger one is
    that workload at one tenth of its vertices and edges).  GPU counts equal
    the conversion identity over the oracle's non-induced P1, P2 and 4-clique
    counts; on the smaller graph also over the closed forms 3-star = sum C(d, 3)
    and tailed triangle = sum t_v (d_v - 2), with t_v from a sparse product."""
    import scipy.sparse as sp
    for n, m, full in [(30_000, 130_000, True), (380_000, 1_650_000, False)]:
        g = J.chung_lu(n, m, gamma=2.3, max_weight=800.0, seed=3)
        _say(f"chung-lu n={g.nv} nnz={g.nnz}")
        G = ctx.upload(g)
        n_c = O.count(g.rowptr, g.col, J.PATTERNS["P1"])
        n_d = O.count(g.rowptr, g.col, J.PATTERNS["P2"])
        d = g.orient()
        n_k = O.count(d.rowptr, d.col, J.PATTERNS["P13"])
        assert G.count("I_K4")[0] == n_k
        assert G.count("I_DIAMOND")[0] == n_d - 6 * n_k
        assert G.count("I_C4")[0] == n_c - n_d + 3 * n_k
        if full:
            deg = g.degrees().astype(np.int64)
            A = sp.csr_matrix((np.ones(g.nnz), g.col, g.rowptr), shape=(g.nv, g.nv))
            tv = np.asarray((A @ A).multiply(A).sum(axis=1)).ravel().astype(np.int64) // 2
            n_s = int((deg * (deg - 1) * (deg - 2) // 6).sum())
            n_t = int((tv * (deg - 2)).sum())
            assert G.count("STAR3")[0] == n_s
            assert G.count("TAILTRI")[0] == n_t
            assert G.count("I_TAILTRI")[0] == n_t - 4 * n_d + 12 * n_k
            assert G.count("I_STAR3")[0] == n_s - n_t + 2 * n_d - 4 * n_k
        G.free()


def test_dag_only_graph_rejects_sym_patterns(ctx):
    g = J.rmat(8, 8, seed=1)
    d = g.orient()
    G = ctx.upload(d, flags=J.DAG)
    assert G.count("TC")[0] == O.count(d.rowptr, d.col, 0)
    with pytest.raises(J.JgpmError) as e:
        G.count("P1")
    assert e.value.status == -7
    G.free()


HEAVY = ["P5", "P7", "P9", "P11", "P12"]


def test_heavy_edge_kernel_thresholds(ctx):
    """P5, P7, P9, P11 and P12 send the edges with more than sym_heavy_min
    common neighbours to the block-per-edge kernel (S in an LDS hash; q(v2)
    by streaming N(v2) or by binary searches, whichever is shorter).  At
    every threshold, from every edge heavy (1) to none (8192), the counts
    equal the oracle's golden counts and quarters on R-MAT 12, the oracle on
    a hub graph, and the closed form on K40 (every edge has 38 common
    neighbours)."""
    ent = GOLD["rmat12_s7"]
    g12 = J.rmat(12, 16, seed=7)
    hub = hub_graph(300, 0.08, seed=11)
    src, dst = np.triu_indices(40, 1)
    k40 = J.CSR.from_edges(40, src, dst)
    want_hub = {p: O.count(hub.rowptr, hub.col, J.PATTERNS[p]) for p in HEAVY}
    want_k40 = {}
    for p in HEAVY:
        k, edges = brute.PATTERNS[J.PATTERNS[p]]
        want_k40[p] = comb(40, k) * int(np.prod(range(1, k + 1))) // brute.automorphisms(k, edges)
    G12, Gh, Gk = ctx.upload(g12), ctx.upload(hub), ctx.upload(k40)
    q = g12.nv // 4
    try:
        for h in [1, 8, 32, 200, 8192]:
            ctx.set_tuning(sym_heavy_min=h)
            assert ctx.tuning()["sym_heavy_min"] == h
            for p in HEAVY:
                _say(f"heavy_min {h} {p}")
                assert G12.count(p)[0] == ent["counts"][p], (h, p)
                assert [G12.count(p, (i * q, (i + 1) * q))[0] for i in range(4)] == ent["quarters"][p], (h, p)
                assert Gh.count(p)[0] == want_hub[p], (h, p)
                assert Gk.count(p)[0] == want_k40[p], (h, p)
    finally:
        ctx.set_tuning(sym_heavy_min=-1)
    assert ctx.tuning()["sym_heavy_min"] == 32
    for G in (G12, Gh, Gk):
        G.free()


@pytest.mark.parametrize("nv_pad", [0, 300_000])
def test_heavy_edge_kernel_wide_shape(ctx, nv_pad):
    """An edge whose endpoints share 1,200 neighbours (two hubs over a sparse
    random graph) runs in the 8192-wide shape of the heavy kernel, the
    others in the narrow one or the wave-per-edge kernel: equal to the
    oracle at every threshold.  Padded with isolated vertices beyond 2^18,
    the kernel keeps S in its hash table instead of a vertex bitmap."""
    rng = np.random.default_rng(3)
    m = 1200
    com = np.arange(2, m + 2)
    iu, ju = np.triu_indices(m, 1)
    keep = rng.random(len(iu)) < 0.01
    src = np.concatenate([[0], np.zeros(m, np.int64), np.ones(m, np.int64), iu[keep] + 2])
    dst = np.concatenate([[1], com, com, ju[keep] + 2])
    g = J.CSR.from_edges(max(m + 2, nv_pad), src, dst)
    G = ctx.upload(g)
    want = {p: O.count(g.rowptr, g.col, J.PATTERNS[p]) for p in HEAVY}
    want2 = {p: O.count(g.rowptr, g.col, J.PATTERNS[p], 0, 2) for p in HEAVY}  # the two hubs
    try:
        for h in [1, 32, 1000, 8192]:
            ctx.set_tuning(sym_heavy_min=h)
            for p in HEAVY:
                _say(f"two hubs nv {g.nv} heavy_min {h} {p}")
                assert G.count(p)[0] == want[p], (h, p)
                assert G.count(p, (0, 2))[0] == want2[p], (h, p)
    finally:
        ctx.set_tuning(sym_heavy_min=-1)
    G.free()


@pytest.mark.parametrize("nv_pad", [0, 1_100_000])
def test_heavy_edge_kernel_huge_shape(ctx, nv_pad):
    """Two hubs sharing 9,000 neighbours: the hub edge's S exceeds the LDS
    list of the 8192-wide shape, so on a graph of at most 2^20 vertices it
    runs in the huge shape (S in the list scratch, a 128 KB vertex bitmap),
    beyond 2^20 vertices in the wave-per-edge kernel.  P5, P7 and P9 equal
    the oracle; P11 and P12 (a minute each in the oracle) equal the
    wave-per-edge kernel's counts, which the other tests hold to the oracle."""
    rng = np.random.default_rng(4)
    m = 9000
    com = np.arange(2, m + 2)
    a = rng.integers(0, m, 40000)
    b = rng.integers(0, m, 40000)
    keep = a != b
    src = np.concatenate([[0], np.zeros(m, np.int64), np.ones(m, np.int64), a[keep] + 2])
    dst = np.concatenate([[1], com, com, b[keep] + 2])
    g = J.CSR.from_edges(max(m + 2, nv_pad), src, dst)
    G = ctx.upload(g)
    want = {}
    for p in HEAVY:
        if p in ("P5", "P7", "P9"):
            want[p] = (O.count(g.rowptr, g.col, J.PATTERNS[p]), O.count(g.rowptr, g.col, J.PATTERNS[p], 0, 2))
        else:
            ctx.set_tuning(sym_heavy_min=8192)
            want[p] = (G.count(p)[0], G.count(p, (0, 2))[0])
    try:
        for h in [1, 32, 8192]:
            ctx.set_tuning(sym_heavy_min=h)
            for p in HEAVY:
                _say(f"9000 common nv {g.nv} heavy_min {h} {p}")
                assert (G.count(p)[0], G.count(p, (0, 2))[0]) == want[p], (h, p)
    finally:
        ctx.set_tuning(sym_heavy_min=-1)
    G.free()


def test_heavy_edge_kernel_rmat16(ctx):
    """On R-MAT 16 (edges with up to ~1,000 common neighbours) the split run
    at the default threshold equals the wave-per-edge run over the whole
    graph and over a window of roots."""
    g = J.rmat(16, 16, seed=5)
    G = ctx.upload(g)
    win = (100, g.nv // 3)
    try:
        for p in HEAVY:
            _say(f"rmat16 heavy {p}")
            ctx.set_tuning(sym_heavy_min=8192)
            off = (G.count(p)[0], G.count(p, win)[0])
            ctx.set_tuning(sym_heavy_min=-1)
            on = (G.count(p)[0], G.count(p, win)[0])
            assert on == off, p
    finally:
        ctx.set_tuning(sym_heavy_min=-1)
    G.free()
