"""General patterns on the GPU (compiled plans, jgpm_count_plan).

Bar: bit-exact.  Every catalogue pattern, compiled from its edge list, gives
the catalogue kernel's total (which the oracle pins); the vertex-induced
4-vertex motifs give the induced-motif counts; random connected patterns
(3..6 vertices, both modes) give the brute-force count on small random
graphs; root ranges add up to the total."""
import random

import numpy as np
import pytest

import jupiter_amd as J
from brute import PATTERNS, adjacency
from plan_brute import count_subgraphs, random_connected_pattern

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)


NAMES = {0: "TC", 17: "STAR3", 18: "TAILTRI", **{i: f"P{i}" for i in range(1, 17)}}


@pytest.fixture(scope="module")
def rmat10(ctx):
    G = ctx.upload(J.rmat(10, 16, seed=21))
    yield G
    G.free()


# The executor walks every partial match, so the patterns with 10^8+
# occurrences per 10^4 edges run on smaller R-MATs (scale 7: P10-P12 have
# 2-7 * 10^7 occurrences; scale 8: P6, P15, P16).
SMALL = {6: 8, 10: 7, 11: 7, 12: 7, 15: 8, 16: 8}


@pytest.mark.parametrize("pid", sorted(NAMES))
def test_catalogue_patterns_equal_kernels(ctx, rmat10, pid):
    k, edges = PATTERNS[pid]
    G = ctx.upload(J.rmat(SMALL[pid], 16, seed=21)) if pid in SMALL else rmat10
    _say(f"plan {NAMES[pid]}")
    got, ms = G.count_plan(J.Plan(k, edges))
    want, _ = G.count(NAMES[pid])
    _say(f"plan {NAMES[pid]}: {got} in {ms:.1f} ms")
    assert got == want
    if G is not rmat10:
        G.free()


@pytest.mark.parametrize("name,pid", [("I_C4", 1), ("I_DIAMOND", 2), ("I_TAILTRI", 18), ("I_STAR3", 17),
                                      ("I_K4", 13)])
def test_induced_motifs(ctx, rmat10, name, pid):
    k, edges = PATTERNS[pid]
    _say(f"induced plan {name}")
    assert rmat10.count_plan(J.Plan(k, edges, induced=True))[0] == rmat10.count(name)[0]


def test_random_patterns_vs_brute_force(ctx):
    rng = random.Random(17)
    for trial in range(20):
        nv = rng.randint(14, 22)
        p = rng.uniform(0.25, 0.5)
        gedges = [(a, b) for a in range(nv) for b in range(a + 1, nv) if rng.random() < p]
        adj = adjacency(nv, gedges)
        src = np.array([a for a, _ in gedges] or [0], dtype=np.int32)
        dst = np.array([b for _, b in gedges] or [0], dtype=np.int32)
        G = ctx.upload(J.CSR.from_edges(nv, src, dst))
        k = rng.randint(3, 5) if nv > 18 else rng.randint(3, 6)
        edges = random_connected_pattern(rng, k, rng.uniform(0.3, 0.8))
        induced = bool(trial % 2)
        _say(f"random pattern k={k} {edges} induced={induced} on n={nv}")
        assert G.count_plan(J.Plan(k, edges, induced))[0] == count_subgraphs(adj, k, edges, induced)
        G.free()


def test_root_ranges_add_up(ctx):
    G = ctx.upload(J.rmat(8, 16, seed=5))
    plan = J.Plan(5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 2)])  # house with a chord
    total, _ = G.count_plan(plan)
    assert total > 0
    cuts = [0, 20, 60, 150, 256]
    assert sum(G.count_plan(plan, (a, b))[0] for a, b in zip(cuts, cuts[1:])) == total
    assert G.count_plan(plan, (5, 5))[0] == 0
    G.free()
