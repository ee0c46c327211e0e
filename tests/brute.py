"""Brute-force subgraph counting for tiny graphs (test-only checker).

count(H, G) = (number of injective maps phi: V(H) -> V(G) sending every edge
of H to an edge of G) / |Aut(H)|, the non-induced subgraph count.  |Aut(H)|
is itself computed by brute force over vertex permutations of H.  This is
independent of the oracle's loop nests and pins them.
"""
import itertools

# Pattern graphs H of each id (vertex labels follow the reference kernels'
# variable names v0..v4 of src/message_passing_gpu_kernels.cuh).
PATTERNS = {
    0: (3, [(0, 1), (0, 2), (1, 2)]),                                   # triangle
    1: (4, [(0, 1), (1, 2), (2, 3), (3, 0)]),                           # 4-cycle
    2: (4, [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3)]),                   # diamond
    3: (5, [(0, 1), (0, 2), (1, 2), (0, 3), (3, 4)]),                   # triangle + 2-path tail
    4: (5, [(0, 1), (0, 2), (1, 2), (1, 3), (3, 4), (4, 0)]),           # house
    5: (5, [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3), (4, 0), (4, 1)]),  # K4 + vertex on edge
    6: (5, [(0, 1), (2, 0), (2, 1), (3, 0), (3, 1), (4, 2), (4, 3)]),   # K_{2,3} + edge
    7: (5, [(0, 1), (0, 2), (0, 3), (0, 4), (4, 1), (1, 2), (2, 3)]),   # gem: hub 0, path 4-1-2-3
    8: (5, [(0, 1), (0, 2), (0, 3), (0, 4), (1, 2), (2, 3), (3, 4), (4, 1)]),  # wheel W4
    9: (5, [(0, 1), (0, 2), (1, 2), (3, 0), (3, 1), (3, 2), (4, 0), (4, 1), (4, 2)]),  # K5 - e
    10: (6, [(0, 1), (0, 2), (0, 3), (0, 4), (0, 5), (4, 1), (1, 2), (2, 3), (3, 5)]),  # 5-fan
    11: (6, [(0, 1), (0, 2), (1, 2), (3, 0), (3, 1), (4, 0), (4, 2), (5, 1), (5, 2)]),  # 3-sun
    12: (6, [(0, 1), (2, 0), (2, 1), (3, 0), (3, 1), (4, 0), (4, 1), (5, 0), (5, 1), (2, 3), (2, 4)]),
    13: (4, [(a, b) for a in range(4) for b in range(a + 1, 4)]),
    14: (5, [(a, b) for a in range(5) for b in range(a + 1, 5)]),
    15: (6, [(a, b) for a in range(6) for b in range(a + 1, 6)]),
    16: (7, [(a, b) for a in range(7) for b in range(a + 1, 7)]),
    # extension ids
    17: (4, [(0, 1), (0, 2), (0, 3)]),                                   # 3-star
    18: (4, [(0, 1), (0, 2), (1, 2), (0, 3)]),                           # tailed triangle
}


def automorphisms(k, edges):
    es = {frozenset(e) for e in edges}
    n = 0
    for perm in itertools.permutations(range(k)):
        if all(frozenset((perm[a], perm[b])) in es for a, b in edges):
            n += 1
    return n


def adjacency(nv, edge_list):
    adj = [set() for _ in range(nv)]
    for u, v in edge_list:
        if u != v:
            adj[u].add(v)
            adj[v].add(u)
    return adj


def count_injective(adj, k, edges):
    """Number of injective maps of H (k vertices, edges) into G."""
    nv = len(adj)
    # order H's vertices so each new vertex has an earlier neighbour when possible
    order = [0]
    while len(order) < k:
        rest = [v for v in range(k) if v not in order]
        rest.sort(key=lambda v: -sum(1 for a, b in edges if (a == v and b in order) or (b == v and a in order)))
        order.append(rest[0])
    back = {v: [u for u in order[:i] if (u, v) in edges or (v, u) in edges] for i, v in enumerate(order)}
    phi = {}
    used = set()

    def rec(i):
        if i == k:
            return 1
        v = order[i]
        nb = back[v]
        if nb:
            cand = set(adj[phi[nb[0]]])
            for u in nb[1:]:
                cand &= adj[phi[u]]
        else:
            cand = range(nv)
        total = 0
        for x in cand:
            if x in used:
                continue
            phi[v] = x
            used.add(x)
            total += rec(i + 1)
            used.discard(x)
        return total

    return rec(0)


def count_pattern(adj, pid):
    k, edges = PATTERNS[pid]
    inj = count_injective(adj, k, edges)
    aut = automorphisms(k, edges)
    assert inj % aut == 0
    return inj // aut
