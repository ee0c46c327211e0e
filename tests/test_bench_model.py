"""bench.py's algorithmic-bytes model of the clique tiers against the
oracle's design-bytes function (each root's list and its members' lists
once), and the tier split against the library's tier widths."""
import random

import numpy as np

import bench
import graphs
import oracle_lib as O
from gen import hub_graph


def test_design_bytes_by_tier_matches_oracle():
    rng = random.Random(5)
    for rp, col in [graphs.random_graph(120, 0.2, rng)[:2], (lambda g: (g.rowptr, g.col))(hub_graph(700, 0.02, 7))]:
        drp, dcol = O.orient(rp, col)
        nv = len(drp) - 1
        deg = np.diff(drp)
        for k in (3, 4, 5):
            for rb, re in ((0, nv), (0, nv // 3), (nv // 2, nv)):
                tiers = bench.design_bytes_by_tier(drp, dcol, rb, re, k)
                skipped = sum(O.design_bytes(drp, dcol, v, v + 1) for v in range(rb, re) if deg[v] < k - 1)
                assert sum(t["bytes"] for t in tiers.values()) + skipped == O.design_bytes(drp, dcol, rb, re)
                assert sum(t["roots"] for t in tiers.values()) == int(((deg[rb:re] >= k - 1)).sum())
        tiers = bench.design_bytes_by_tier(drp, dcol, 0, nv, 3)
        assert tiers["clique_lds1024"]["roots"] == int(((deg > 512) & (deg <= 1024)).sum())
