"""Brute-force references for general patterns (test-only checkers).

count_subgraphs(adj, k, edges, induced): the number of subgraphs of G
isomorphic to H (non-induced), or of vertex subsets inducing H (induced),
as injective maps / |Aut(H)|.  plan_matches(...) enumerates injective maps
in a compiled plan's level order and keeps those satisfying its
restrictions (v_a < v_b): with a correct plan the two numbers agree."""
import itertools

from brute import automorphisms


def _maps(adj, k, edges, induced, order=None):
    """All injective maps of H into G (level order `order`), as tuples phi[v]."""
    nv = len(adj)
    es = {frozenset(e) for e in edges}
    order = order or list(range(k))
    phi = [None] * k

    def ok(v, x):
        for u in range(k):
            if phi[u] is None or u == v:
                continue
            e = frozenset((u, v)) in es
            a = x in adj[phi[u]]
            if e and not a:
                return False
            if induced and a and not e:
                return False
        return True

    def rec(i):
        if i == k:
            yield tuple(phi)
            return
        v = order[i]
        used = set(x for x in phi if x is not None)
        for x in range(nv):
            if x in used or not ok(v, x):
                continue
            phi[v] = x
            yield from rec(i + 1)
            phi[v] = None

    yield from rec(0)


def count_subgraphs(adj, k, edges, induced=False):
    n = sum(1 for _ in _maps(adj, k, edges, induced))
    aut = automorphisms(k, edges)
    assert n % aut == 0
    return n // aut


def plan_matches(adj, k, edges, induced, order, restrictions):
    """Injective maps satisfying the plan's restrictions; `order[l]` is the
    pattern vertex of level l, restrictions are (a, b) over levels."""
    n = 0
    for phi in _maps(adj, k, edges, induced, order):
        lv = [phi[order[l]] for l in range(k)]
        if all(lv[a] < lv[b] for a, b in restrictions):
            n += 1
    return n


def random_connected_pattern(rng, k, p):
    while True:
        edges = [(a, b) for a, b in itertools.combinations(range(k), 2) if rng.random() < p]
        seen, stack = {0}, [0]
        while stack:
            u = stack.pop()
            for a, b in edges:
                for x, y in ((a, b), (b, a)):
                    if x == u and y not in seen:
                        seen.add(y)
                        stack.append(y)
        if len(seen) == k and edges:
            return edges
