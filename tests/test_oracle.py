"""Pins the CPU oracle (the parity checker) without the reference:
brute-force subgraph counts, complete-graph closed forms, matrix closed
forms, invariants, and the set primitives against Python sets."""
import random
from math import comb

import numpy as np
import pytest

import brute
import graphs
import oracle_lib as O

DAG_IDS = (0, 13, 14, 15, 16)
SYM_IDS = tuple(range(1, 13))
# |Aut(H)| per id as tabulated for the reference's pattern catalogue
AUT = {0: 6, 1: 8, 2: 4, 3: 2, 4: 2, 5: 4, 6: 4, 7: 2, 8: 8, 9: 12, 10: 2, 11: 6, 12: 4,
       13: 24, 14: 120, 15: 720, 16: 5040}


def ocount(rp, col, pid, mode=1, rb=0, re=None):
    if pid in DAG_IDS:
        drp, dcol = O.orient(rp, col)
        return O.count(drp, dcol, pid, rb, re, mode=mode)
    return O.count(rp, col, pid, rb, re, mode=mode)


@pytest.mark.parametrize("seed", range(8))
def test_oracle_matches_brute_force(seed):
    rng = random.Random(seed)
    nv = rng.randint(5, 10)
    rp, col, adj = graphs.random_graph(nv, rng.uniform(0.25, 0.85), rng)
    for pid in DAG_IDS + SYM_IDS:
        bf = brute.count_pattern(adj, pid)
        assert ocount(rp, col, pid, mode=0) == bf, pid
        assert ocount(rp, col, pid, mode=1) == bf, pid


def test_automorphism_table():
    for pid, aut in AUT.items():
        k, edges = brute.PATTERNS[pid]
        assert brute.automorphisms(k, edges) == aut, pid


@pytest.mark.parametrize("n", range(1, 10))
def test_complete_graph_closed_form(n):
    rp, col, _ = graphs.complete_graph(n)
    for pid in DAG_IDS + SYM_IDS:
        k, _e = brute.PATTERNS[pid]
        expect = comb(n, k) * int(np.prod(range(1, k + 1))) // AUT[pid]
        assert ocount(rp, col, pid) == expect, (n, pid)


def test_complete_bipartite():
    # K_{m,n}: only the 4-cycle survives, C(m,2) C(n,2) of them
    m, n = 5, 7
    rp, col, _ = graphs.csr_from_edges(m + n, [(a, m + b) for a in range(m) for b in range(n)])
    for pid in DAG_IDS + SYM_IDS:
        expect = comb(m, 2) * comb(n, 2) if pid == 1 else 0
        assert ocount(rp, col, pid) == expect, pid


@pytest.mark.parametrize("seed", range(3))
def test_matrix_closed_forms(seed):
    rng = random.Random(100 + seed)
    nv = 70
    rp, col, adj = graphs.random_graph(nv, 0.15, rng)
    A = np.zeros((nv, nv), dtype=np.int64)
    for u in range(nv):
        A[u, list(adj[u])] = 1
    A2 = A @ A
    assert ocount(rp, col, 0) == np.trace(A2 @ A) // 6
    # 4-cycles: half the sum over unordered pairs of C(codegree, 2)
    iu = np.triu_indices(nv, 1)
    assert ocount(rp, col, 1) == sum(comb(int(c), 2) for c in A2[iu]) // 2
    # diamonds: sum over edges of C(t_e, 2)
    t = A2 * A
    assert ocount(rp, col, 2) == sum(comb(int(t[u, v]), 2) for u in range(nv) for v in range(u + 1, nv) if A[u, v])


def test_relabelling_invariance():
    rng = random.Random(5)
    nv = 40
    rp, col, adj = graphs.random_graph(nv, 0.2, rng)
    perm = list(range(nv))
    rng.shuffle(perm)
    rp2, col2, _ = graphs.relabel(nv, adj, perm)
    for pid in DAG_IDS + SYM_IDS:
        assert ocount(rp, col, pid) == ocount(rp2, col2, pid), pid


def test_quarter_sums():
    rng = random.Random(9)
    nv = 48
    rp, col, _ = graphs.random_graph(nv, 0.2, rng)
    q = nv // 4
    for pid in DAG_IDS + SYM_IDS:
        total = ocount(rp, col, pid)
        assert sum(ocount(rp, col, pid, rb=i * q, re=(i + 1) * q) for i in range(4)) == total, pid


def test_literal_and_adjusted_modes_agree():
    rng = random.Random(11)
    rp, col, _ = graphs.random_graph(60, 0.12, rng)
    for pid in SYM_IDS:
        assert ocount(rp, col, pid, mode=0) == ocount(rp, col, pid, mode=1), pid


@pytest.mark.parametrize("seed", range(20))
def test_set_primitives(seed):
    rng = np.random.default_rng(seed)
    na = int(rng.integers(0, 200))
    nb = int(rng.integers(0, 3)) if seed % 3 == 0 else int(rng.integers(0, 200))  # gallop regime too
    a = np.unique(rng.integers(0, 300, na)).astype(np.int32)
    b = np.unique(rng.integers(0, 300, nb)).astype(np.int32)
    for upper in (2**31 - 1, 150, 0):
        ref = sorted(x for x in set(a.tolist()) & set(b.tolist()) if x < upper)
        assert O.isect(a, b, upper).tolist() == ref
        assert O.isect_num(a, b, upper) == len(ref)
        # set difference: A − B, with A − ∅ = A
        assert O.diff(a, b, upper).tolist() == sorted(x for x in set(a.tolist()) - set(b.tolist()) if x < upper)
