"""jgpm_count_noalloc (the clique counts without device allocation, scratch
in a static per-GPU array) against jgpm_count through the C ABI: the same
counts on consecutive calls, the second on a smaller input, including roots
of the 2048-wide tier (rows in the static scratch) and the other tiers."""
import ctypes

import numpy as np
import pytest

import jupiter_amd as J
from gen import hub_graph

pytestmark = pytest.mark.gpu


def noalloc(G, pattern, roots=None):
    rb, re = (0, -1) if roots is None else roots
    L = J.lib()
    f = L.jgpm_count_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64)]
    f.restype = ctypes.c_int
    out = ctypes.c_uint64(0)
    st = f(G._h, J.pattern_id(pattern), rb, re, ctypes.byref(out))
    if st:
        raise J.JgpmError(st, "count_noalloc")
    return int(out.value)


def complete(n):
    src, dst = np.triu_indices(n, 1)
    return J.CSR.from_edges(n, src, dst)


def test_noalloc_matches_count(ctx):
    big = ctx.upload(complete(1100))        # 75 roots wider than 1024: the 2048-wide tier
    hub = ctx.upload(hub_graph(1500, 0.012, seed=1500))
    mid = ctx.upload(J.rmat(12, 16, seed=7))
    small = ctx.upload(J.rmat(9, 16, seed=1))
    try:
        for name, G, pats in (("K1100", big, ("TC", "P13")), ("hub1500", hub, ("TC", "P13", "P14", "P15")),
                              ("rmat12", mid, ("TC", "P13", "P14", "I_K4")), ("rmat9", small, ("TC", "P13", "P14"))):
            for p in pats:
                print(f"[case] noalloc {name} {p}", flush=True)
                want = G.count(p)[0]
                assert noalloc(G, p) == want, (name, p)
                # a second call in a row on a smaller input: a root range
                nv = G.info()["nv"]
                r = (0, max(1, nv // 3))
                assert noalloc(G, p, r) == G.count(p, r)[0], (name, p, r)
        # consecutive calls moving from the largest input to the smallest
        print("[case] noalloc K1100 then rmat9", flush=True)
        assert noalloc(big, "P13") == big.count("P13")[0]
        assert noalloc(small, "P13") == small.count("P13")[0]
        assert noalloc(small, "TC", (0, 0)) == 0
        with pytest.raises(J.JgpmError) as e:
            noalloc(small, "P1")
        assert e.value.status == -2
        with pytest.raises(J.JgpmError) as e:
            noalloc(small, "P13", (5, 3))
        assert e.value.status == -1
    finally:
        for G in (big, hub, mid, small):
            G.free()
