"""Serving loop: concurrently submitted queries of different patterns and
graphs give the same uint64 count as alone (tolerance 0); admission by the
HBM budget rejects what cannot fit."""
import threading

import numpy as np
import pytest

import jupiter_amd as J

pytestmark = pytest.mark.gpu


def test_mixed_concurrent_requests_equal_alone(ctx):
    graphs = [ctx.upload(J.rmat(11, 16, seed=s)) for s in (1, 2)]
    graphs.append(ctx.upload(J.barabasi_albert(4000, 6, seed=3)))
    graphs.append(ctx.upload(J.erdos_renyi(3000, 16.0, seed=4)))
    rng = np.random.default_rng(0)
    pats = ["TC"] + [f"P{i}" for i in range(1, 17)]
    reqs = []
    for i in range(60):
        G = graphs[int(rng.integers(0, len(graphs)))]
        p = pats[int(rng.integers(0, len(pats)))]
        nv = G.info()["nv"]
        if i % 3 == 0:
            a = int(rng.integers(0, nv))
            b = int(rng.integers(a, nv + 1))
            reqs.append((G, p, (a, b)))
        else:
            reqs.append((G, p, None))
    alone = [G.count(p, r)[0] for G, p, r in reqs]
    results = [None] * len(reqs)
    with J.Server(ctx, slice_roots=512) as srv:
        def worker(ids):
            for k in ids:
                t = srv.submit(reqs[k][0], reqs[k][1], reqs[k][2])
                results[k] = srv.wait(t)
        ths = [threading.Thread(target=worker, args=(list(range(j, len(reqs), 4)),)) for j in range(4)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
    assert [r[0] for r in results] == alone
    assert all(r[1] > 0 for r in results)
    for G in graphs:
        G.free()


def test_heavy_and_light_requests_mixed(ctx):
    """Requests whose work exceeds one slice run whole with the single-query
    kernels between the batched launches of the light ones; both kinds in one
    stream of submissions give their counts alone."""
    big = ctx.upload(J.rmat(16, 16, seed=5))      # P3, P5, P13, P14, P1 whole-graph: heavy
    small = ctx.upload(J.rmat(10, 16, seed=6))    # everything light
    reqs = [(big, "P3"), (small, "P2"), (big, "P13"), (small, "P5"), (big, "P1"), (small, "TC"),
            (big, "P14"), (small, "P9"), (big, "TC"), (big, "P1", (100, 5000)), (small, "P13")]
    alone = [r[0].count(r[1], r[2] if len(r) > 2 else None)[0] for r in reqs]
    with J.Server(ctx) as srv:
        tickets = [srv.submit(r[0], r[1], r[2] if len(r) > 2 else None) for r in reqs]
        got = [srv.wait(t)[0] for t in tickets]
    assert got == alone
    big.free()
    small.free()


def test_budget_and_pattern_rejection(ctx):
    G = ctx.upload(J.rmat(10, 16, seed=1))
    with J.Server(ctx, hbm_budget_bytes=1 << 16) as srv:
        with pytest.raises(J.JgpmError) as e:
            srv.submit(G, "P3")
        assert e.value.status == -6
        with pytest.raises(J.JgpmError) as e:
            srv.submit(G, "I_C4")
        assert e.value.status == -2
        t = srv.submit(G, "TC", (5, 5))
        assert srv.wait(t) == (0, 0.0)
        # a result is handed out once, then the ticket is forgotten
        with pytest.raises(J.JgpmError) as e:
            srv.wait(t)
        assert e.value.status == -1
        with pytest.raises(J.JgpmError) as e:
            srv.poll(t)
        assert e.value.status == -1
    G.free()


def test_heavy_lanes_on_fresh_graphs(ctx):
    """Heavy requests run on lanes, each a context of its own, so several are
    on the device at once.  On graphs whose per-graph arrays are not built yet
    (COO sources, edge triangles, neighbour-degree sums, the degree-order
    relabelling of a whole-graph 4-cycle count, the ids of a root wider than
    2048) the scheduler builds them before a lane runs the query; every count
    equals the single-query count computed afterwards, with four lanes and
    with one."""
    from gen import hub_graph
    csrs = [J.rmat(15, 16, seed=s) for s in (7, 8)] + [hub_graph(2600, 0.006, seed=26)]
    spec = [(0, "P1", None), (1, "P3", None), (0, "P7", None), (1, "P13", None), (0, "P14", None),
            (1, "P1", None), (0, "P3", None), (2, "TC", None), (2, "P13", None), (1, "P4", (30000, 32768)),
            (0, "P2", None), (1, "P1", (50, 20000)), (1, "P11", None), (0, "P2", (100, 9000))]
    got = {}
    for lanes in (4, 1):
        # fresh uploads each time: the scheduler builds every per-graph array
        gs = [ctx.upload(c) for c in csrs]
        try:
            print(f"[case] serving {len(spec)} heavy requests on fresh graphs, {lanes} lane(s)", flush=True)
            with J.Server(ctx, heavy_lanes=lanes) as srv:
                tickets = [srv.submit(gs[i], p, r) for i, p, r in spec]
                got[lanes] = [srv.wait(t)[0] for t in tickets]
        finally:
            for G in gs:
                G.free()
    # alone: one query per fresh upload, so no array is shared with the runs above
    alone = []
    for i, p, r in spec:
        G = ctx.upload(csrs[i])
        try:
            alone.append(G.count(p, r)[0])
        finally:
            G.free()
    assert got[4] == alone
    assert got[1] == alone
