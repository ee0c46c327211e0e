"""Small-graph helpers for the tests (pure numpy; independent of libjgpm)."""
import numpy as np


def csr_from_edges(nv, edges):
    """Symmetric CSR with ascending, duplicate-free lists and no self loops."""
    adj = [set() for _ in range(nv)]
    for u, v in edges:
        u, v = int(u), int(v)
        if u != v:
            adj[u].add(v)
            adj[v].add(u)
    rowptr = np.zeros(nv + 1, dtype=np.int64)
    for v in range(nv):
        rowptr[v + 1] = rowptr[v] + len(adj[v])
    col = np.zeros(max(int(rowptr[-1]), 1), dtype=np.int32)
    for v in range(nv):
        col[rowptr[v]:rowptr[v + 1]] = sorted(adj[v])
    return rowptr, col[: rowptr[-1]], adj


def random_graph(nv, p, rng):
    edges = [(u, v) for u in range(nv) for v in range(u + 1, nv) if rng.random() < p]
    return csr_from_edges(nv, edges)


def complete_graph(n):
    return csr_from_edges(n, [(u, v) for u in range(n) for v in range(u + 1, n)])


def relabel(nv, adj, perm):
    edges = [(perm[u], perm[v]) for u in range(nv) for v in adj[u] if u < v]
    return csr_from_edges(nv, edges)
