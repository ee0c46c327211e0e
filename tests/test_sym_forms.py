"""CPU checks of the algebraic forms the symmetric kernels compute, against
the oracle (the restated reference nests):

* whole-graph P3 as a sum over triangles a < b < c of three oriented terms
  (a, b, c), (b, a, c), (c, a, b), the form of the triangle kernel
  (p3_tri_kernel), with a planted mistake (the orientation (a, c, b) in place
  of (c, a, b)) that must change the total on the same graphs;
* P5 per edge as (|S| - 2) |E(S)|, each edge of S counted once from its
  smaller end, the form of the heavy-edge kernel (sym_heavy_kernel<5>).
"""
import numpy as np
import pytest

import jupiter_amd as J
import oracle_lib as O


def p3_by_triangles(g, plant=False):
    rp = np.asarray(g.rowptr)
    ci = np.asarray(g.col)
    nv = len(rp) - 1
    nbr = [set(ci[rp[v]:rp[v + 1]].tolist()) for v in range(nv)]
    d = np.diff(rp).astype(np.int64)
    nsum = np.array([int(d[list(nbr[v])].sum()) if nbr[v] else 0 for v in range(nv)], dtype=np.int64)

    def t(u, v):
        return len(nbr[u] & nbr[v])

    def term(v0, v1, v2):
        # X(v0, v1) - (d(v2) - 2) - (t(v0, v2) - 1), X = D0 - d1 - d0 + 1 - t(v0, v1)
        return (int(nsum[v0]) - int(d[v1]) - int(d[v0]) + 1 - t(v0, v1)) - (int(d[v2]) - 2) - (t(v0, v2) - 1)

    tot = 0
    for a in range(nv):
        for b in nbr[a]:
            if b <= a:
                continue
            for c in nbr[a] & nbr[b]:
                if c <= b:
                    continue
                tot += term(a, b, c) + term(b, a, c) + (term(a, c, b) if plant else term(c, a, b))
    return tot


@pytest.mark.parametrize("scale,seed", [(7, 2), (9, 4)])
def test_p3_triangle_form_equals_oracle(scale, seed):
    g = J.rmat(scale, 8, seed=seed)
    want = O.count(g.rowptr, g.col, J.PATTERNS["P3"])
    assert p3_by_triangles(g) == want
    assert p3_by_triangles(g, plant=True) != want


def p5_by_edges_of_s(g, plant=False):
    """P5 as the heavy-edge kernel computes it: per edge (v0, v1), v1 < v0,
    (|S| - 2) |E(S)| with S = N(v0) ∩ N(v1), each edge of S counted once from
    its smaller end (the planted mistake counts it from both ends)."""
    rp = np.asarray(g.rowptr)
    ci = np.asarray(g.col)
    nv = len(rp) - 1
    nbr = [set(ci[rp[v]:rp[v + 1]].tolist()) for v in range(nv)]
    tot = 0
    for v0 in range(nv):
        for v1 in nbr[v0]:
            if v1 >= v0:
                continue
            s = nbr[v0] & nbr[v1]
            n = len(s)
            if n < 3:
                continue
            e = sum(1 for v2 in s for x in nbr[v2] if x in s and (plant or x > v2))
            tot += (n - 2) * e
    return tot


@pytest.mark.parametrize("scale,seed", [(7, 2), (9, 4)])
def test_p5_edges_of_s_form_equals_oracle(scale, seed):
    g = J.rmat(scale, 8, seed=seed)
    want = O.count(g.rowptr, g.col, J.PATTERNS["P5"])
    assert p5_by_edges_of_s(g) == want
    assert p5_by_edges_of_s(g, plant=True) != want
