"""Seeded test graphs shared by the CPU and GPU tests."""
import numpy as np

import jupiter_amd as J


def hub_graph(core, p, seed):
    """A root (vertex 0) adjacent to `core` vertices that each have degree
    >= core (pendant leaves pad them), plus a sparse random graph among the
    core vertices.  In the degree DAG vertex 0 then has out-degree `core`,
    which exercises the wide-row tiers, while clique counts stay small."""
    rng = np.random.default_rng(seed)
    cores = np.arange(1, core + 1, dtype=np.int64)
    src, dst = [], []
    src.append(np.zeros(core, dtype=np.int64)); dst.append(cores)
    iu, ju = np.triu_indices(core, 1)
    keep = rng.random(len(iu)) < p
    src.append(iu[keep] + 1); dst.append(ju[keep] + 1)
    deg = np.ones(core, dtype=np.int64)
    np.add.at(deg, iu[keep], 1)
    np.add.at(deg, ju[keep], 1)
    need = np.maximum(0, core - deg)
    nleaf = int(need.sum())
    leaf_ids = core + 1 + np.arange(nleaf, dtype=np.int64)
    src.append(np.repeat(cores, need)); dst.append(leaf_ids)
    nv = core + 1 + nleaf
    return J.CSR.from_edges(nv, np.concatenate(src), np.concatenate(dst))
