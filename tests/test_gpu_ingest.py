"""On-device ingest: graphs built on the GPU (from an edge list, or sampled
by the device R-MAT generator) are array-for-array identical to the host
construction (symmetric CSR and degree DAG), and count the same.

Bar: bit-exact arrays and counts."""
import numpy as np
import pytest

import jupiter_amd as J

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _same(a, b):
    assert a.nv == b.nv and a.nnz == b.nnz
    assert np.array_equal(a.rowptr, b.rowptr)
    assert np.array_equal(a.col, b.col)


@pytest.mark.parametrize("scale,ef,seed", [(8, 8, 1), (12, 16, 7), (16, 16, 3)])
def test_device_rmat_equals_host(ctx, scale, ef, seed):
    _say(f"device rmat {scale} ef{ef} seed{seed}")
    host = J.rmat(scale, ef, seed=seed)
    G = ctx.rmat_graph(scale, ef, seed=seed)
    _same(G.download(), host)
    _same(G.download(dag=True), host.orient())
    H = ctx.upload(host)
    assert G.info() == H.info()
    for p in ("TC", "P13", "P2"):
        assert G.count(p)[0] == H.count(p)[0]
    G.free()
    H.free()


def test_device_edges_with_loops_and_duplicates(ctx):
    rng = np.random.default_rng(4)
    nv = 3000
    src = rng.integers(0, nv, 40000).astype(np.int32)
    dst = rng.integers(0, nv, 40000).astype(np.int32)
    src[:500] = dst[:500]                 # self loops
    src = np.concatenate([src, dst[:2000]])   # reversed duplicates
    dst = np.concatenate([dst, src[:2000]])
    _say("device edge list with loops and duplicates")
    host = J.CSR.from_edges(nv, src, dst)
    G = ctx.graph_from_edges(nv, src, dst)
    _same(G.download(), host)
    _same(G.download(dag=True), host.orient())
    H = ctx.upload(host)
    assert G.count("P14")[0] == H.count("P14")[0]
    G.free()
    H.free()


def test_device_edges_edge_cases(ctx):
    G = ctx.graph_from_edges(5, [], [])
    assert G.info()["sym_nnz"] == 0 and G.count("TC")[0] == 0
    G.free()
    G = ctx.graph_from_edges(1, [0], [0])  # one vertex, one loop
    assert G.info()["sym_nnz"] == 0
    G.free()
    with pytest.raises(J.JgpmError) as e:
        ctx.graph_from_edges(4, [0, 5], [1, 2])  # id out of range
    assert e.value.status == -1


def test_file_load_orients_on_the_device(ctx, tmp_path):
    """jgpm_graph_load maps the reference's files, checks them in place and
    builds the degree DAG on the device: the arrays equal the host
    construction, counts equal an upload, and a DAG-only load keeps the DAG
    alone."""
    host = J.rmat(14, 16, seed=4)
    prefix = str(tmp_path / "g")
    host.write(prefix)
    _say("file load, symmetric + DAG")
    G = ctx.load(prefix)
    U = ctx.upload(host)
    try:
        _same(G.download(), host)
        _same(G.download(dag=True), host.orient())
        for p in ("TC", "P13", "P1", "P2"):
            assert G.count(p)[0] == U.count(p)[0], p
    finally:
        G.free()
        U.free()
    _say("file load, DAG only")
    D = ctx.load(prefix, dag_only=True)
    try:
        _same(D.download(dag=True), host.orient())
        assert D.info()["sym_nnz"] == -1
        with pytest.raises(J.JgpmError) as e:
            D.count("P2")
        assert e.value.status == -7
        assert D.count("P14")[0] > 0
    finally:
        D.free()
