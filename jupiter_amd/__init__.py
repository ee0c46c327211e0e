"""jupiter_amd — MI355X-native graph pattern counting (Python host mirror).

Thin ctypes layer over the C ABI of ``include/jgpm.h`` (``libjgpm.so``, HIP
kernels for gfx950).  It mirrors the reference's operator surface:

* ``Graph``: a resident CSR graph (reference ``Graph(prefix, use_dag)``,
  common/graph.cc:8-124 plus ``NvsGraphGPU::init``, include/graph_gpu_nvs.h:258-304);
* ``Graph.count("P13", roots=(b, e))``: one pattern query over a root range
  (the reference's ``SubgraphSolver`` dispatch, src/jupiter_gpm.cu:391-908);
* pattern names ``P1``..``P16`` as in the reference, plus ``TC`` and the
  induced 4-vertex motif ids.

There is no CPU fallback: if ``libjgpm.so`` is missing or no GPU is present the
calls raise.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libjgpm.so")

PATTERNS = {
    "TC": 0,
    **{f"P{i}": i for i in range(1, 17)},
    "STAR3": 17,
    "TAILTRI": 18,
    "I_C4": 20,
    "I_DIAMOND": 21,
    "I_TAILTRI": 22,
    "I_STAR3": 23,
    "I_K4": 24,
}
PATTERN_NAMES = {v: k for k, v in PATTERNS.items()}
DAG_PATTERNS = {0, 13, 14, 15, 16, 24}

SYM = 1
DAG = 2

STATUS = {
    0: "ok", -1: "invalid argument", -2: "unknown pattern id", -3: "out of memory", -4: "HIP runtime error",
    -5: "graph file missing or malformed", -6: "over budget / unsupported size", -7: "needs symmetric graph",
    -8: "busy", -9: "not ready", -10: "bad state", -11: "no HIP device",
}


class JgpmError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        super().__init__(f"{what}: {STATUS.get(status, status)} ({status})")


_lib = None


class _Tuning(ctypes.Structure):
    _fields_ = [("small_max_deg", ctypes.c_int32), ("lds_max_deg", ctypes.c_int32),
                ("gallop_ratio", ctypes.c_int32), ("waves_per_block", ctypes.c_int32),
                ("blocks_per_cu", ctypes.c_int32), ("batch_hub_deg", ctypes.c_int32),
                ("batch_root_chunk", ctypes.c_int32), ("batch_edge_chunk", ctypes.c_int32),
                ("batch_blocks_per_cu", ctypes.c_int32), ("sym_heavy_min", ctypes.c_int32)]


class _Query(ctypes.Structure):
    _fields_ = [("graph", ctypes.c_void_p), ("pattern", ctypes.c_int32), ("_pad", ctypes.c_int32),
                ("root_begin", ctypes.c_int64), ("root_end", ctypes.c_int64)]


class _ServeCfg(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("max_inflight", ctypes.c_int32),
                ("slice_roots", ctypes.c_int32), ("heavy_lanes", ctypes.c_int32), ("reserved", ctypes.c_int32)]


def lib():
    """Load libjgpm.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m jupiter_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    P, I32, I64, U64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64
    PP = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "jgpm_init": ([ctypes.c_int, PP], ctypes.c_int),
        "jgpm_shutdown": ([P], ctypes.c_int),
        "jgpm_strerror": ([ctypes.c_int], ctypes.c_char_p),
        "jgpm_device_info": ([P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                              ctypes.POINTER(I64), ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
        "jgpm_graph_upload": ([P, P, P, I32, I64, ctypes.c_uint32, PP], ctypes.c_int),
        "jgpm_graph_load": ([P, ctypes.c_char_p, ctypes.c_int, PP], ctypes.c_int),
        "jgpm_graph_free": ([P], ctypes.c_int),
        "jgpm_graph_info": ([P, ctypes.POINTER(I32), ctypes.POINTER(I64), ctypes.POINTER(I64),
                             ctypes.POINTER(I32), ctypes.POINTER(I32)], ctypes.c_int),
        "jgpm_graph_bytes": ([P, ctypes.POINTER(I64)], ctypes.c_int),
        "jgpm_count": ([P, ctypes.c_int, I64, I64, ctypes.POINTER(U64), ctypes.POINTER(ctypes.c_double)],
                       ctypes.c_int),
        "jgpm_work_bytes": ([P, ctypes.c_int, I64, I64, ctypes.POINTER(U64)], ctypes.c_int),
        "jgpm_get_tuning": ([P, ctypes.POINTER(_Tuning)], ctypes.c_int),
        "jgpm_last_profile": ([P, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_double)],
                              ctypes.c_int),
        "jgpm_work_bytes_split": ([P, ctypes.c_int, I64, I64, ctypes.POINTER(U64), ctypes.c_int], ctypes.c_int),
        "jgpm_set_tuning": ([P, ctypes.POINTER(_Tuning)], ctypes.c_int),
        "jgpm_count_batch": ([P, ctypes.POINTER(_Query), ctypes.c_int, ctypes.POINTER(U64),
                              ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "jgpm_plan_compile": ([I32, I32, P, P, ctypes.c_int, PP], ctypes.c_int),
        "jgpm_plan_describe": ([P, ctypes.POINTER(I32), ctypes.POINTER(I64), ctypes.POINTER(I32), ctypes.POINTER(I32),
                                ctypes.POINTER(I32)], ctypes.c_int),
        "jgpm_count_plan": ([P, P, I64, I64, ctypes.POINTER(U64), ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "jgpm_plan_free": ([P], ctypes.c_int),
        "jgpm_graph_from_edges": ([P, I32, I64, P, P, PP], ctypes.c_int),
        "jgpm_graph_rmat": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, U64,
                             PP], ctypes.c_int),
        "jgpm_graph_download": ([P, ctypes.c_int, PP], ctypes.c_int),
        "jgpm_batch_create": ([P, ctypes.POINTER(_Query), ctypes.c_int, PP], ctypes.c_int),
        "jgpm_batch_run": ([P, ctypes.POINTER(U64), ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "jgpm_batch_free": ([P], ctypes.c_int),
        "jgpm_autotune_batch": ([P, ctypes.POINTER(_Query), ctypes.c_int, ctypes.c_int, ctypes.POINTER(_Tuning),
                                 ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "jgpm_serve_start": ([P, ctypes.POINTER(_ServeCfg), PP], ctypes.c_int),
        "jgpm_serve_submit": ([P, P, ctypes.c_int, I64, I64, ctypes.POINTER(U64)], ctypes.c_int),
        "jgpm_serve_poll": ([P, U64, ctypes.POINTER(U64), ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "jgpm_serve_wait": ([P, U64, ctypes.POINTER(U64), ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "jgpm_serve_stop": ([P], ctypes.c_int),
        "jgpm_gen_rmat": ([ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, U64, PP],
                          ctypes.c_int),
        "jgpm_gen_chung_lu": ([I32, I64, ctypes.c_double, ctypes.c_double, U64, PP], ctypes.c_int),
        "jgpm_gen_er": ([I32, ctypes.c_double, U64, PP], ctypes.c_int),
        "jgpm_gen_ba": ([I32, I32, U64, PP], ctypes.c_int),
        "jgpm_csr_from_edges": ([I32, I64, P, P, PP], ctypes.c_int),
        "jgpm_csr_orient": ([P, PP], ctypes.c_int),
        "jgpm_csr_info": ([P, ctypes.POINTER(I32), ctypes.POINTER(I64), ctypes.POINTER(I32)], ctypes.c_int),
        "jgpm_csr_data": ([P, PP, PP], ctypes.c_int),
        "jgpm_csr_write": ([P, ctypes.c_char_p], ctypes.c_int),
        "jgpm_csr_read": ([ctypes.c_char_p, PP], ctypes.c_int),
        "jgpm_csr_free": ([P], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(st, what):
    if st != 0:
        raise JgpmError(st, what)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def pattern_id(p):
    if isinstance(p, str):
        if p not in PATTERNS:
            raise JgpmError(-2, f"pattern {p!r}")
        return PATTERNS[p]
    return int(p)


# ---------------------------------------------------------------------------
# Host CSR


class CSR:
    """Host CSR (int64 rowptr, int32 ascending colidx) held in numpy arrays."""

    def __init__(self, rowptr, col):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int32)

    @property
    def nv(self):
        return len(self.rowptr) - 1

    @property
    def nnz(self):
        return int(self.rowptr[-1])

    def degrees(self):
        return np.diff(self.rowptr)

    @staticmethod
    def _from_handle(h):
        L = lib()
        nv, nnz, md = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32()
        _check(L.jgpm_csr_info(h, ctypes.byref(nv), ctypes.byref(nnz), ctypes.byref(md)), "csr_info")
        rp, ci = ctypes.c_void_p(), ctypes.c_void_p()
        _check(L.jgpm_csr_data(h, ctypes.byref(rp), ctypes.byref(ci)), "csr_data")
        r = np.ctypeslib.as_array(ctypes.cast(rp, ctypes.POINTER(ctypes.c_int64)), (nv.value + 1,)).copy()
        if nnz.value:
            c = np.ctypeslib.as_array(ctypes.cast(ci, ctypes.POINTER(ctypes.c_int32)), (nnz.value,)).copy()
        else:
            c = np.zeros(0, dtype=np.int32)
        L.jgpm_csr_free(h)
        return CSR(r, c)

    def _handle(self):
        """Temporary library-side copy (through the edge list) of this CSR."""
        src = np.repeat(np.arange(self.nv, dtype=np.int32), np.diff(self.rowptr))
        h = ctypes.c_void_p()
        _check(lib().jgpm_csr_from_edges(self.nv, self.nnz, _ptr(src), _ptr(self.col), ctypes.byref(h)),
               "csr_from_edges")
        return h

    def orient(self):
        """Degree-ordered DAG (common/graph.cc:229-230)."""
        h = self._handle()
        d = ctypes.c_void_p()
        st = lib().jgpm_csr_orient(h, ctypes.byref(d))
        lib().jgpm_csr_free(h)
        _check(st, "csr_orient")
        return CSR._from_handle(d)

    def write(self, prefix):
        h = self._handle()
        st = lib().jgpm_csr_write(h, prefix.encode())
        lib().jgpm_csr_free(h)
        _check(st, "csr_write")

    @staticmethod
    def read(prefix):
        h = ctypes.c_void_p()
        _check(lib().jgpm_csr_read(prefix.encode(), ctypes.byref(h)), "csr_read")
        return CSR._from_handle(h)

    @staticmethod
    def from_edges(nv, src, dst):
        src = np.ascontiguousarray(src, dtype=np.int32)
        dst = np.ascontiguousarray(dst, dtype=np.int32)
        h = ctypes.c_void_p()
        _check(lib().jgpm_csr_from_edges(nv, len(src), _ptr(src), _ptr(dst), ctypes.byref(h)), "csr_from_edges")
        return CSR._from_handle(h)


def rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1):
    h = ctypes.c_void_p()
    _check(lib().jgpm_gen_rmat(scale, edge_factor, a, b, c, seed, ctypes.byref(h)), "gen_rmat")
    return CSR._from_handle(h)


def chung_lu(n, target_edges, gamma=2.3, max_weight=800.0, seed=1):
    h = ctypes.c_void_p()
    _check(lib().jgpm_gen_chung_lu(n, target_edges, gamma, max_weight, seed, ctypes.byref(h)), "gen_chung_lu")
    return CSR._from_handle(h)


def erdos_renyi(n, mean_degree, seed=1):
    h = ctypes.c_void_p()
    _check(lib().jgpm_gen_er(n, mean_degree, seed, ctypes.byref(h)), "gen_er")
    return CSR._from_handle(h)


def barabasi_albert(n, m, seed=1):
    h = ctypes.c_void_p()
    _check(lib().jgpm_gen_ba(n, m, seed, ctypes.byref(h)), "gen_ba")
    return CSR._from_handle(h)


# ---------------------------------------------------------------------------
# Device context and resident graphs


class Context:
    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        _check(lib().jgpm_init(device, ctypes.byref(self._h)), "init")

    def close(self):
        if self._h:
            lib().jgpm_shutdown(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self):
        cu, lds, hbm = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        name = ctypes.create_string_buffer(128)
        _check(lib().jgpm_device_info(self._h, ctypes.byref(cu), ctypes.byref(lds), ctypes.byref(hbm), name, 128),
               "device_info")
        return {"num_cu": cu.value, "lds_bytes": lds.value, "hbm_bytes": hbm.value, "arch": name.value.decode()}

    def tuning(self):
        t = _Tuning()
        _check(lib().jgpm_get_tuning(self._h, ctypes.byref(t)), "get_tuning")
        return {f: getattr(t, f) for f, _ in _Tuning._fields_}

    def set_tuning(self, **kw):
        t = _Tuning(**{f: int(kw.get(f, 0)) for f, _ in _Tuning._fields_})
        _check(lib().jgpm_set_tuning(self._h, ctypes.byref(t)), "set_tuning")

    def last_profile(self):
        """[(kernel name, ms)] of the last count on this context, in launch order."""
        names = (ctypes.c_char_p * 16)()
        ms = (ctypes.c_double * 16)()
        n = lib().jgpm_last_profile(self._h, 16, names, ms)
        if n < 0:
            raise JgpmError(n, "last_profile")
        return [(names[i].decode(), ms[i]) for i in range(min(n, 16))]

    def upload(self, csr, flags=SYM):
        return Graph(self, csr, flags)

    def graph_from_edges(self, nv, src, dst):
        """Build a resident graph on the GPU from an undirected edge list (int32 arrays)."""
        src = np.ascontiguousarray(src, dtype=np.int32)
        dst = np.ascontiguousarray(dst, dtype=np.int32)
        h = ctypes.c_void_p()
        _check(lib().jgpm_graph_from_edges(self._h, nv, len(src), _ptr(src), _ptr(dst), ctypes.byref(h)),
               "graph_from_edges")
        return Graph._wrap(self, h)

    def rmat_graph(self, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1):
        """Seeded R-MAT sampled and built on the GPU (the same graph as rmat(...) uploaded)."""
        h = ctypes.c_void_p()
        _check(lib().jgpm_graph_rmat(self._h, scale, edge_factor, a, b, c, seed, ctypes.byref(h)), "graph_rmat")
        return Graph._wrap(self, h)

    def load(self, prefix, dag_only=False):
        g = Graph.__new__(Graph)
        g.ctx = self
        g._h = ctypes.c_void_p()
        _check(lib().jgpm_graph_load(self._h, prefix.encode(), int(dag_only), ctypes.byref(g._h)), "graph_load")
        return g

    @staticmethod
    def _queries(queries):
        arr = (_Query * len(queries))()
        for i, (g, p, rb, re) in enumerate(queries):
            arr[i].graph = g._h
            arr[i].pattern = pattern_id(p)
            arr[i].root_begin = rb
            arr[i].root_end = -1 if re is None else re
        return arr

    def count_batch(self, queries):
        """queries: list of (Graph, pattern, root_begin, root_end). Returns (counts, kernel_ms)."""
        arr = self._queries(queries)
        out = (ctypes.c_uint64 * len(queries))()
        ms = ctypes.c_double()
        _check(lib().jgpm_count_batch(self._h, arr, len(queries), out, ctypes.byref(ms)), "count_batch")
        return [int(x) for x in out], ms.value

    def prepare_batch(self, queries):
        """A Batch: the queries' launches captured once in a HIP graph, replayed by Batch.run()."""
        return Batch(self, queries)

    def autotune_batch(self, queries, reps=2):
        """Tune the batched-launch knobs on these queries; installs and returns (tuning, best_ms)."""
        arr = self._queries(queries)
        t = _Tuning()
        ms = ctypes.c_double()
        _check(lib().jgpm_autotune_batch(self._h, arr, len(queries), reps, ctypes.byref(t), ctypes.byref(ms)),
               "autotune_batch")
        return {f: getattr(t, f) for f, _ in _Tuning._fields_}, ms.value


class Plan:
    """A compiled general pattern (jgpm_plan_*): k vertices, edges [(a, b), ...]."""

    def __init__(self, k, edges, induced=False):
        eu = np.ascontiguousarray([a for a, _ in edges], dtype=np.int32)
        ev = np.ascontiguousarray([b for _, b in edges], dtype=np.int32)
        self.k = k
        self._h = ctypes.c_void_p()
        _check(lib().jgpm_plan_compile(k, len(edges), _ptr(eu), _ptr(ev), int(induced), ctypes.byref(self._h)),
               "plan_compile")

    def describe(self):
        """(matching order, |Aut|, restrictions [(a, b)] meaning v_a < v_b over levels)."""
        order = (ctypes.c_int32 * self.k)()
        aut = ctypes.c_int64()
        ra = (ctypes.c_int32 * 64)()
        rb = (ctypes.c_int32 * 64)()
        n = ctypes.c_int32(64)
        _check(lib().jgpm_plan_describe(self._h, order, ctypes.byref(aut), ra, rb, ctypes.byref(n)), "plan_describe")
        return list(order), aut.value, [(ra[i], rb[i]) for i in range(n.value)]

    def free(self):
        if self._h:
            lib().jgpm_plan_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Batch:
    """Prepared batched count (jgpm_batch_*): run() replays the captured launches."""

    def __init__(self, ctx, queries):
        self.ctx = ctx
        self.n = len(queries)
        self._h = ctypes.c_void_p()
        arr = Context._queries(queries)
        _check(lib().jgpm_batch_create(ctx._h, arr, self.n, ctypes.byref(self._h)), "batch_create")

    def run(self):
        out = (ctypes.c_uint64 * self.n)()
        ms = ctypes.c_double()
        _check(lib().jgpm_batch_run(self._h, out, ctypes.byref(ms)), "batch_run")
        return [int(x) for x in out], ms.value

    def free(self):
        if self._h:
            _check(lib().jgpm_batch_free(self._h), "batch_free")
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()


class Graph:
    """A CSR graph resident in device memory."""

    def __init__(self, ctx, csr, flags=SYM):
        self.ctx = ctx
        self._h = ctypes.c_void_p()
        _check(lib().jgpm_graph_upload(ctx._h, _ptr(csr.rowptr), _ptr(csr.col), csr.nv, csr.nnz, flags,
                                       ctypes.byref(self._h)), "graph_upload")

    @classmethod
    def _wrap(cls, ctx, handle):
        g = cls.__new__(cls)
        g.ctx = ctx
        g._h = handle
        return g

    def count_plan(self, plan, roots=None):
        """Count a compiled general pattern; roots select the level-0 vertices. Returns (count, kernel_ms)."""
        rb, re = (0, -1) if roots is None else roots
        out = ctypes.c_uint64()
        ms = ctypes.c_double()
        _check(lib().jgpm_count_plan(self._h, plan._h, rb, re, ctypes.byref(out), ctypes.byref(ms)), "count_plan")
        return int(out.value), ms.value

    def download(self, dag=False):
        """The resident symmetric (or DAG) CSR copied back to the host as a CSR."""
        h = ctypes.c_void_p()
        _check(lib().jgpm_graph_download(self._h, int(dag), ctypes.byref(h)), "graph_download")
        return CSR._from_handle(h)

    def free(self):
        if self._h:
            _check(lib().jgpm_graph_free(self._h), "graph_free")
            self._h = ctypes.c_void_p()

    def info(self):
        nv, sn, dn, sm, dm = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        _check(lib().jgpm_graph_info(self._h, ctypes.byref(nv), ctypes.byref(sn), ctypes.byref(dn),
                                     ctypes.byref(sm), ctypes.byref(dm)), "graph_info")
        return {"nv": nv.value, "sym_nnz": sn.value, "dag_nnz": dn.value, "sym_max_deg": sm.value,
                "dag_max_deg": dm.value}

    def count(self, pattern, roots=None):
        """Count `pattern` over roots [b, e) (default: all). Returns (count, kernel_ms)."""
        rb, re = (0, -1) if roots is None else roots
        out = ctypes.c_uint64()
        ms = ctypes.c_double()
        _check(lib().jgpm_count(self._h, pattern_id(pattern), rb, re, ctypes.byref(out), ctypes.byref(ms)),
               f"count {pattern}")
        return int(out.value), ms.value

    def work_bytes_split(self, pattern, roots=None):
        """Work-model bytes per kernel launch of the query (order of last_profile)."""
        rb, re = (0, -1) if roots is None else roots
        out = (ctypes.c_uint64 * 16)()
        n = lib().jgpm_work_bytes_split(self._h, pattern_id(pattern), rb, re, out, 16)
        if n < 0:
            raise JgpmError(n, "work_bytes_split")
        return [int(out[i]) for i in range(min(n, 16))]

    def work_bytes(self, pattern, roots=None):
        rb, re = (0, -1) if roots is None else roots
        out = ctypes.c_uint64()
        _check(lib().jgpm_work_bytes(self._h, pattern_id(pattern), rb, re, ctypes.byref(out)), "work_bytes")
        return int(out.value)


class Server:
    """Continuous-batching server over graphs resident on one context
    (jgpm_serve_*).  submit() returns a ticket; wait()/poll() give
    (count, latency_us)."""

    def __init__(self, ctx, hbm_budget_bytes=0, max_inflight=0, slice_roots=0, heavy_lanes=0):
        cfg = _ServeCfg(hbm_budget_bytes, max_inflight, slice_roots, heavy_lanes, 0)
        self._h = ctypes.c_void_p()
        _check(lib().jgpm_serve_start(ctx._h, ctypes.byref(cfg), ctypes.byref(self._h)), "serve_start")

    def submit(self, graph, pattern, roots=None):
        rb, re = (0, -1) if roots is None else roots
        t = ctypes.c_uint64()
        _check(lib().jgpm_serve_submit(self._h, graph._h, pattern_id(pattern), rb, re, ctypes.byref(t)),
               "serve_submit")
        return t.value

    def poll(self, ticket):
        c, us = ctypes.c_uint64(), ctypes.c_double()
        st = lib().jgpm_serve_poll(self._h, ticket, ctypes.byref(c), ctypes.byref(us))
        if st == -9:
            return None
        _check(st, "serve_poll")
        return int(c.value), us.value

    def wait(self, ticket):
        c, us = ctypes.c_uint64(), ctypes.c_double()
        _check(lib().jgpm_serve_wait(self._h, ticket, ctypes.byref(c), ctypes.byref(us)), "serve_wait")
        return int(c.value), us.value

    def stop(self):
        if self._h:
            lib().jgpm_serve_stop(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.stop()
