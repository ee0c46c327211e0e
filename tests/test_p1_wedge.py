"""The 4-cycle (P1) wedge identity the GPU kernel uses, checked on the CPU.

The reference's P1 value of a directed edge (v0, v1) with v1 < v0 is
  sum over v2 in N(v0), v2 < v1, of |N(v1) ∩ N(v2) ∩ [0, v0)|
(src/message_passing_gpu_kernels.cuh:104-186).  The kernel computes, per
root v0, sum over x < v0 of C(c_x, 2) with c_x = |{v in N(v0): v < v0,
x in N(v)}|.  These tests compare the two per root (so every root range
gives the reference's partial count), and the total against the closed form
1/2 sum_{u<w} C(codeg(u, w), 2) (every 4-cycle has two diagonals).
"""
import itertools

import numpy as np
import pytest

from graphs import complete_graph, random_graph


def nest_per_root(nv, adj):
    out = np.zeros(nv, dtype=np.int64)
    for v0 in range(nv):
        n0 = sorted(adj[v0])
        for v1 in n0:
            if v1 >= v0:
                continue
            for v2 in n0:
                if v2 >= v1:
                    break
                out[v0] += sum(1 for x in adj[v1] & adj[v2] if x < v0)
    return out


def wedge_per_root(nv, adj):
    out = np.zeros(nv, dtype=np.int64)
    for v0 in range(nv):
        c = {}
        for v in adj[v0]:
            if v >= v0:
                continue
            for x in adj[v]:
                if x < v0:
                    c[x] = c.get(x, 0) + 1
        out[v0] = sum(k * (k - 1) // 2 for k in c.values())
    return out


def c4_closed_form(nv, adj):
    s = 0
    for u, w in itertools.combinations(range(nv), 2):
        k = len(adj[u] & adj[w])
        s += k * (k - 1) // 2
    return s // 2


@pytest.mark.parametrize("seed,nv,p", [(1, 12, 0.5), (2, 20, 0.3), (3, 30, 0.2), (4, 16, 0.9)])
def test_wedge_equals_nest_per_root(seed, nv, p):
    rng = np.random.default_rng(seed)
    _, _, adj = random_graph(nv, p, rng)
    a = nest_per_root(nv, adj)
    b = wedge_per_root(nv, adj)
    assert np.array_equal(a, b)
    assert int(b.sum()) == c4_closed_form(nv, adj)


@pytest.mark.parametrize("n", [4, 5, 8, 11])
def test_wedge_complete_graph(n):
    _, _, adj = complete_graph(n)
    # K_n has C(n,4) * 3 four-cycles (4!/|Aut(C4)| = 3 per 4-set)
    assert int(wedge_per_root(n, adj).sum()) == 3 * (n * (n - 1) * (n - 2) * (n - 3) // 24)
