"""The pattern -> plan compiler (host only): for every pattern of the
catalogue and random connected patterns, the compiled matching order and
symmetry-breaking restrictions select exactly one injective map per
occurrence (brute force on small random graphs), non-induced and induced;
disconnected patterns are rejected."""
import random

import pytest

import jupiter_amd as J
from brute import PATTERNS, adjacency, automorphisms
from plan_brute import count_subgraphs, plan_matches, random_connected_pattern


def _graph(rng, nv, p):
    edges = [(a, b) for a in range(nv) for b in range(a + 1, nv) if rng.random() < p]
    return adjacency(nv, edges)


CASES = [(k, edges) for pid, (k, edges) in sorted(PATTERNS.items()) if k <= 6]


@pytest.mark.parametrize("k,edges", CASES)
@pytest.mark.parametrize("induced", [False, True])
def test_catalogue_plans(k, edges, induced):
    plan = J.Plan(k, edges, induced)
    order, aut, restr = plan.describe()
    assert sorted(order) == list(range(k))
    assert aut == automorphisms(k, edges)
    rng = random.Random(k * 100 + len(edges) + induced)
    nonzero = False
    for nv, p in ((7, 0.6), (8, 0.45), (8, 0.85)):  # the dense graph holds cliques up to K6
        adj = _graph(rng, nv, p)
        ref = count_subgraphs(adj, k, edges, induced)
        assert plan_matches(adj, k, edges, induced, order, restr) == ref
        nonzero = nonzero or ref > 0
    assert nonzero or induced


def test_random_patterns():
    rng = random.Random(9)
    for trial in range(25):
        k = rng.randint(3, 6)
        edges = random_connected_pattern(rng, k, rng.uniform(0.3, 0.8))
        induced = bool(trial % 2)
        order, aut, restr = J.Plan(k, edges, induced).describe()
        adj = _graph(rng, 7, 0.55)
        assert plan_matches(adj, k, edges, induced, order, restr) == count_subgraphs(adj, k, edges, induced), edges


def test_rejects():
    with pytest.raises(J.JgpmError) as e:
        J.Plan(4, [(0, 1), (2, 3)])  # not connected
    assert e.value.status == -2
    with pytest.raises(J.JgpmError):
        J.Plan(8, [(0, 1)])  # too many vertices
    with pytest.raises(J.JgpmError):
        J.Plan(3, [(0, 0)])  # loop
