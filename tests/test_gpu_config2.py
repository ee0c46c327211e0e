"""The benchmark workload at its own size: R-MAT scale 22 (edge factor 16,
a/b/c/d = 0.57/0.19/0.19/0.05, seed 1), the graph sampled, symmetrised,
deduplicated and oriented on the device.

* The downloaded symmetric and DAG CSR equal the host build (SHA-256 of
  rowptr and colidx frozen in tests/golden/rmat22.json).
* TC, P13 and P14 over root windows of the first quarter [0, 2^20) that
  together hold roots of every clique tier equal the oracle's counts frozen
  by tests/golden/make_golden.py --rmat22; TC and P13 over the whole quarter
  equal the oracle's (frozen the same way).
* TC, P13 and P14 over the four quarters (the reference's per-GPU root
  ranges, src/jupiter_gpm.cu:153-162) add up to the whole-graph count.
"""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rmat22.json")))


def sha(c):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(c.rowptr, dtype=np.int64).tobytes())
    h.update(np.ascontiguousarray(c.col, dtype=np.int32).tobytes())
    return h.hexdigest()


@pytest.fixture(scope="module")
def g22(ctx):
    print("[case] R-MAT 22 built on the device", flush=True)
    G = ctx.rmat_graph(22, 16, 0.57, 0.19, 0.19, seed=1)
    yield G
    G.free()


def test_device_graph_equals_host_build(g22):
    info = g22.info()
    assert info["nv"] == GOLD["nv"] and info["sym_nnz"] == GOLD["sym_nnz"] and info["dag_nnz"] == GOLD["dag_nnz"]
    assert info["dag_max_deg"] == GOLD["dag_max_deg"]
    print("[case] download symmetric CSR", flush=True)
    assert sha(g22.download(dag=False)) == GOLD["sym_sha256"]
    print("[case] download DAG CSR", flush=True)
    assert sha(g22.download(dag=True)) == GOLD["dag_sha256"]


@pytest.mark.parametrize("w", range(len(GOLD["windows"])))
def test_root_windows_against_oracle(g22, w):
    win = GOLD["windows"][w]
    a, b = win["roots"]
    for p in ("TC", "P13", "P14"):
        print(f"[case] {p} roots [{a},{b})", flush=True)
        assert g22.count(p, (a, b))[0] == win[p], (p, a, b)


def test_windows_cover_every_tier():
    tiers = {k: sum(w["tier_roots"][k] for w in GOLD["windows"]) for k in ("64", "128", "256", "512", "1024")}
    assert all(v > 0 for v in tiers.values()), tiers


def test_quarter_against_oracle(g22):
    q = GOLD["nv"] // 4
    for p, want in sorted(GOLD.get("quarter0", {}).items()):
        print(f"[case] {p} roots [0,{q})", flush=True)
        assert g22.count(p, (0, q))[0] == want, p


def test_quarters_add_up(g22):
    q = GOLD["nv"] // 4
    for p in ("TC", "P13", "P14"):
        print(f"[case] {p} quarters", flush=True)
        parts = [g22.count(p, (i * q, (i + 1) * q))[0] for i in range(4)]
        assert sum(parts) == g22.count(p)[0], (p, parts)
