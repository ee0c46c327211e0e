"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: it checks the HIP path, it is never the
path under test.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = ctypes.CDLL(_LIB)
        P = ctypes.c_void_p
        L.oracle_count.argtypes = [P, P, ctypes.c_int32, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                   ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_count_roots.argtypes = [P, P, ctypes.c_int32, ctypes.c_int, P, ctypes.c_int64,
                                         ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_work_bytes.argtypes = [P, P, ctypes.c_int32, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                        ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_design_bytes.argtypes = [P, P, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64,
                                          ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_vertex_sums.argtypes = [P, P, ctypes.c_int32, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                         ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_orient.argtypes = [P, P, ctypes.c_int32, P, P]
        L.oracle_orient.restype = ctypes.c_int64
        L.oracle_isect.argtypes = [P, ctypes.c_int32, P, ctypes.c_int32, ctypes.c_int32, P]
        L.oracle_isect.restype = ctypes.c_int32
        L.oracle_isect_num.argtypes = [P, ctypes.c_int32, P, ctypes.c_int32, ctypes.c_int32]
        L.oracle_isect_num.restype = ctypes.c_uint64
        L.oracle_diff.argtypes = [P, ctypes.c_int32, P, ctypes.c_int32, ctypes.c_int32, P]
        L.oracle_diff.restype = ctypes.c_int32
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def count(rowptr, col, pid, rb=0, re=None, mode=1, threads=0):
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    nv = len(rowptr) - 1
    if re is None:
        re = nv
    out = ctypes.c_uint64(0)
    rc = lib().oracle_count(_p(rowptr), _p(col), nv, pid, rb, re, mode, threads, ctypes.byref(out))
    assert rc == 0, rc
    return out.value


def count_roots(rowptr, col, pid, roots, mode=1, threads=0):
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    roots = np.ascontiguousarray(roots, dtype=np.int32)
    out = ctypes.c_uint64(0)
    rc = lib().oracle_count_roots(_p(rowptr), _p(col), len(rowptr) - 1, pid, _p(roots), len(roots),
                                  mode, threads, ctypes.byref(out))
    assert rc == 0, rc
    return out.value


def work_bytes(rowptr, col, pid, rb=0, re=None, threads=0):
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    nv = len(rowptr) - 1
    if re is None:
        re = nv
    out = ctypes.c_uint64(0)
    rc = lib().oracle_work_bytes(_p(rowptr), _p(col), nv, pid, rb, re, threads, ctypes.byref(out))
    assert rc == 0, rc
    return out.value


def vertex_sums(rowptr, col, mode, rb=0, re=None):
    """mode 0: 3-stars sum C(d,3); mode 1: tailed triangles sum t_v (d_v - 2)."""
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    nv = len(rowptr) - 1
    if re is None:
        re = nv
    out = ctypes.c_uint64(0)
    rc = lib().oracle_vertex_sums(_p(rowptr), _p(col), nv, mode, rb, re, ctypes.byref(out))
    assert rc == 0, rc
    return out.value


def design_bytes(rowptr, col, rb=0, re=None):
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    nv = len(rowptr) - 1
    if re is None:
        re = nv
    out = ctypes.c_uint64(0)
    rc = lib().oracle_design_bytes(_p(rowptr), _p(col), nv, rb, re, ctypes.byref(out))
    assert rc == 0, rc
    return out.value


def orient(rowptr, col):
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    nv = len(rowptr) - 1
    nnz = lib().oracle_orient(_p(rowptr), _p(col), nv, None, None)
    orp = np.zeros(nv + 1, dtype=np.int64)
    oc = np.zeros(max(nnz, 1), dtype=np.int32)
    lib().oracle_orient(_p(rowptr), _p(col), nv, _p(orp), _p(oc))
    return orp, oc[:nnz]


def isect(a, b, upper=2**31 - 1):
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    out = np.zeros(max(1, min(len(a), len(b))), dtype=np.int32)
    n = lib().oracle_isect(_p(a), len(a), _p(b), len(b), upper, _p(out))
    return out[:n]


def isect_num(a, b, upper=2**31 - 1):
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    return lib().oracle_isect_num(_p(a), len(a), _p(b), len(b), upper)


def diff(a, b, upper=2**31 - 1):
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    out = np.zeros(max(1, len(a)), dtype=np.int32)
    n = lib().oracle_diff(_p(a), len(a), _p(b), len(b), upper, _p(out))
    return out[:n]
