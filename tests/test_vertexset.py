"""The drop-in VertexSet (include/VertexSet.h) against the reference's own
VertexSet (common/VertexSet.cc, compiled from the reference tree by
oracle/Makefile into oracle/_ref/), and the oracle's set primitives against
both.  Skipped when the reference tree is not present."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libref_vs.so")
MINE = os.path.join(ROOT, "oracle", "_build", "libmy_vs.so")


def _load(path):
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    L.vs_set_op.argtypes = [ctypes.c_int, P, ctypes.c_int32, ctypes.c_int32, P, ctypes.c_int32, ctypes.c_int32,
                            ctypes.c_int32, P]
    L.vs_set_op.restype = ctypes.c_int32
    L.vs_count_op.argtypes = [ctypes.c_int, P, ctypes.c_int32, ctypes.c_int32, P, ctypes.c_int32, ctypes.c_int32,
                              ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    L.vs_count_op.restype = ctypes.c_uint64
    L.vs_set_max_degree.argtypes = [ctypes.c_int32]
    L.vs_set_max_degree(4096)
    return L


@pytest.fixture(scope="module")
def libs():
    if not os.path.isdir("/root/reference/include"):
        pytest.skip("reference tree not present")
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "vs"])
    return _load(REF), _load(MINE)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.mark.parametrize("seed", range(30))
def test_vertexset_matches_reference(libs, seed):
    ref, mine = libs
    rng = np.random.default_rng(seed)
    na, nb = int(rng.integers(0, 120)), int(rng.integers(0, 120))
    a = np.unique(rng.integers(0, 200, na)).astype(np.int32)
    b = np.unique(rng.integers(0, 200, nb)).astype(np.int32)
    avid = int(rng.integers(0, 200))
    bvid = int(a[rng.integers(0, len(a))]) if len(a) and seed % 2 else int(rng.integers(0, 200))
    for upper in (2**31 - 1, 100, 0, int(rng.integers(0, 200))):
        for op in range(5):
            o1 = np.zeros(256, dtype=np.int32)
            o2 = np.zeros(256, dtype=np.int32)
            n1 = ref.vs_set_op(op, _p(a), len(a), avid, _p(b), len(b), bvid, upper, _p(o1))
            n2 = mine.vs_set_op(op, _p(a), len(a), avid, _p(b), len(b), bvid, upper, _p(o2))
            assert n1 == n2 and o1[:n1].tolist() == o2[:n2].tolist(), (op, upper)
        anc, anc2 = int(rng.integers(0, 200)), (int(b[0]) if len(b) else 7)
        for op in range(7):
            c1 = ref.vs_count_op(op, _p(a), len(a), avid, _p(b), len(b), bvid, upper, anc, anc2)
            c2 = mine.vs_count_op(op, _p(a), len(a), avid, _p(b), len(b), bvid, upper, anc, anc2)
            assert c1 == c2, (op, upper)
        # the oracle's intersection primitives follow the same merge
        o1 = np.zeros(256, dtype=np.int32)
        n1 = ref.vs_set_op(1, _p(a), len(a), avid, _p(b), len(b), bvid, upper, _p(o1))
        assert O.isect(a, b, upper).tolist() == o1[:n1].tolist()
        assert O.isect_num(a, b, upper) == ref.vs_count_op(1, _p(a), len(a), avid, _p(b), len(b), bvid, upper, 0, 0)
