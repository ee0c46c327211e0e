"""Benchmark of the pattern-counting path on one MI355X.

The timed workload: 4-clique (P13) and 5-clique (P14) counting over one
quarter of the roots of a seeded R-MAT graph (scale 22, edge factor 16,
a/b/c/d = 0.57/0.19/0.19/0.05, degree-ordered DAG, int32 CSR, uint64
counts).  A step is one P13 query plus one P14 query over roots [0, 2^20),
each a plain per-call launch that returns its exact count to the host.
`value` is queries/s over the timed steps (inputs resident in HBM before
timing); the counts are checked against the oracle's frozen counts.

Also on the same JSON line:
  roofline      the dominant clique tier (largest device time, HIP events on
                the library's stream; for the split 1024-wide tier its build
                and count launches): algorithmic bytes of the bitmap design
                (every root reads its own list and each member's list once,
                16 + 4 d+ bytes per list) / its average time per query, against
                the 8 TB/s HBM peak; `traffic` is the measured HBM bytes of the
                same launches (rocprofv3 FETCH_SIZE + WRITE_SIZE summarised in
                profiles/), reported only while the kernel sources are the ones
                that were measured;
  other_configs the triangle count on R-MAT 14 (launch-bound latency) and
                the vertex-induced 4-vertex motifs on the 3.8M-vertex
                Chung-Lu graph, each checked against frozen oracle counts;
  batched       the 2,048-graph batch (half Erdos-Renyi, half Barabasi-
                Albert, one pattern each from {TC, P13, P1, P2}): autotuned,
                captured and replayed vs one plain launch per query;
  serving       Poisson request streams over 32 resident R-MAT graphs of
                scales 18-24 (uniform graph, uniform pattern id P1..P16 within
                the stated per-pattern caps), continuous batching vs the
                plain one-query-per-launch path, equal counts required;
  cpu_baseline  the CPU oracle (clean-room OpenMP restatement of the
                reference nests) on a stratified sample of the same roots,
                all host threads and one thread, scaled to the full query with
                its standard error.
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-serve]
       python bench.py --workload batch|serve     (the batched / serving
                                                    workloads on their own)
"""
import argparse
import hashlib
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import jupiter_amd as J  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "queries/s served vs plain per-call kernel (bit-exact counts); p50/p99; GB/s"
SCALE, EF, SEED = 22, 16, 1
QUERIES = ("P13", "P14")
GOLDEN = os.path.join(ROOT, "tests", "golden")
TIER_NAMES = ("clique_wave64", "clique_lds128", "clique_lds256", "clique_lds512", "clique_lds1024", "clique_glb2048")
TIER_MAX = (64, 128, 256, 512, 1024, 2048)
CLIQUE_K = {"TC": 3, "P13": 4, "P14": 5, "P15": 6, "P16": 7}
# kernel sources whose change makes a committed traffic measurement stale
CLIQUE_SOURCES = ("jupiter_amd/csrc/clique_kernels.hip", "jupiter_amd/csrc/clique_dev.hpp")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def source_digest(paths=CLIQUE_SOURCES):
    h = hashlib.sha256()
    for p in paths:
        with open(os.path.join(ROOT, p), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def design_bytes_by_tier(rowptr, col, rb, re, k):
    """Algorithmic bytes of the bitmap clique engine per tier over the roots
    [rb, re) of a DAG: a root of out-degree d in [k-1, 2048] reads its list and
    each member's list once, 16 + 4 d+ bytes per list (two int64 row pointers
    and the int32 ids); roots with fewer than k-1 members hold no k-clique and
    are skipped by the kernels."""
    rowptr = np.asarray(rowptr, dtype=np.int64)
    d = np.diff(rowptr)
    c = 16 + 4 * d
    # sum over members of c, per root (empty rows give 0)
    cs = np.concatenate([[0], np.cumsum(c[np.asarray(col, dtype=np.int64)])])
    member = cs[rowptr[1:]] - cs[rowptr[:-1]]
    roots = np.arange(rb, re)
    dr = d[rb:re]
    per_root = c[rb:re] + member[rb:re]
    out = {}
    lo = k - 1
    for name, hi in zip(TIER_NAMES, TIER_MAX):
        m = (dr >= lo) & (dr <= hi)
        out[name] = {"roots": int(m.sum()), "bytes": int(per_root[m].sum())}
        lo = hi + 1
    del roots
    return out


def measured_traffic(launch_names, pattern):
    """HBM bytes (read + written) of one pass of the given launches on the
    default workload, from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary in
    profiles/ (scaled by the plain-copy calibration).  None when a launch is
    not measured or the clique sources changed since the measurement."""
    path = os.path.join(ROOT, "profiles", "clique_p13_p14_rmat22_q0_traffic.json")
    if not os.path.exists(path):
        return None
    t = json.load(open(path))
    if t.get("sources_sha256_16") != source_digest():
        return None
    tbl = t.get("kernels", {})
    keys = [f"{n} ({pattern})" for n in launch_names]
    if any(k not in tbl for k in keys):
        return None
    return int(sum(tbl[k]["read_bytes"] + tbl[k]["write_bytes"] for k in keys))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(dag, roots_end, seed=7, per_stratum=(256, 64, 32, 16, 8), batches=4):
    """Oracle (P13 + P14) on a stratified random sample of the roots
    [0, roots_end).  Strata are the clique tiers by DAG out-degree.  One
    thread: each stratum's sample is timed in `batches` random batches; the
    stratum's time is (its roots) x (mean time per sampled root), with the
    standard error from the spread of the batch means.  All threads: the
    whole sample in one dynamically scheduled call gives the parallel speedup
    over the one-thread time of the same roots, which scales the one-thread
    estimate."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    rng = np.random.default_rng(seed)
    d = np.diff(dag.rowptr)[:roots_end]
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    bounds = [(0, 64), (65, 128), (129, 256), (257, 512), (513, 2048)]
    strata = []
    for (lo, hi), m in zip(bounds, per_stratum):
        ids = np.nonzero((d >= lo) & (d <= hi))[0]
        if len(ids) == 0:
            continue
        take = np.sort(rng.choice(ids, size=min(m, len(ids)), replace=False)).astype(np.int32)
        strata.append((len(ids), take))

    def timed(roots, nthreads):
        t0 = time.perf_counter()
        for p in QUERIES:
            O.count_roots(dag.rowptr, dag.col, J.PATTERNS[p], roots, mode=1, threads=nthreads)
        return time.perf_counter() - t0

    est1, var1, secs1 = 0.0, 0.0, 0.0
    for n_s, take in strata:
        means = []
        for part in np.array_split(rng.permutation(take), batches):
            dt = timed(np.sort(part), 1)
            secs1 += dt
            means.append(dt / max(1, len(part)))
        means = np.array(means)
        est1 += n_s * means.mean()
        var1 += (n_s ** 2) * means.var(ddof=1) / len(means)
    # parallel speedup on 16,384 roots of the two light strata, whose many
    # small roots keep every thread busy as the full query's 10^6 roots do
    # (the few sampled hub roots alone cannot: the slowest of them bounds a
    # parallel call)
    pool = np.nonzero(d <= 128)[0]
    light = np.sort(rng.choice(pool, size=min(16384, len(pool)), replace=False)).astype(np.int32)
    light1 = timed(light, 1)
    wall = timed(light, threads)
    speedup = light1 / wall
    se1 = float(np.sqrt(var1))
    q = len(QUERIES)
    nroots = sum(len(t) for _, t in strata)
    return {"value": round(q * speedup / est1, 6), "unit": "queries/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"stratified by clique tier: {nroots} of {roots_end} roots "
                      f"({', '.join(str(len(t)) + '/' + str(n) for n, t in strata)}), P13+P14, one thread "
                      f"{secs1:.1f}s; all {threads} threads: speedup {speedup:.1f} measured on the "
                      f"{len(light)} light-stratum roots ({light1:.2f}s -> {wall:.3f}s)",
            "est_s_per_step": round(est1 / speedup, 1), "std_err_s": round(se1 / speedup, 1),
            "parallel_speedup": round(speedup, 2),
            "one_thread": {"value": round(q / est1, 7), "est_s_per_step": round(est1, 1), "std_err_s": round(se1, 1)}}


def splitmix(seed, i):
    x = (seed * 0x9E3779B97F4A7C15 + i * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    x ^= x >> 31
    return (x * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF


def config4_specs(n_graphs=2048, max_n=262144):
    """The batched workload's graphs: (kind, n, parameter, seed, pattern),
    half Erdos-Renyi (mean degree U[8,64]), half Barabasi-Albert (m U{4..16}),
    n log-uniform in [1k, max_n], pattern uniform from {TC, P13, P1, P2}."""
    rng = np.random.default_rng(SEED)
    out = []
    for i in range(n_graphs):
        n = int(np.exp(rng.uniform(np.log(1000), np.log(max_n))))
        sd = splitmix(SEED, i)
        if i % 2 == 0:
            spec = ("ER", n, float(rng.uniform(8, 64)), sd)
        else:
            spec = ("BA", n, int(rng.integers(4, 17)), sd)
        out.append(spec + (["TC", "P13", "P1", "P2"][int(rng.integers(0, 4))],))
    return out


def config4_graph(spec):
    kind, n, par, sd, _ = spec
    return J.erdos_renyi(n, par, seed=sd) if kind == "ER" else J.barabasi_albert(n, par, seed=sd)


def workload_batch(args, ctx=None):
    """Many independent graphs, each paired with one pattern from
    {TC, P13, P1, P2}: the batch autotuned, captured once in a HIP graph and
    replayed, vs one plain launch per query.  The default is the full
    configuration (2,048 graphs, n up to 262,144); its counts are checked
    against the plain launches and a sample against the oracle's frozen
    counts."""
    specs = config4_specs(args.batch_graphs, args.batch_max_n)
    t0 = time.time()
    edges = 0
    own = ctx is None
    if own:
        ctx = J.Context(0)
    Gs = []
    for i, spec in enumerate(specs):
        c = config4_graph(spec)
        edges += c.nnz // 2
        Gs.append(ctx.upload(c))  # resident; the host copy is dropped
        if (i + 1) % 128 == 0:
            log(f"generated and uploaded {i + 1} graphs, {edges} edges ({time.time() - t0:.1f}s)")
    gen_s = time.time() - t0
    log(f"generated {len(Gs)} graphs, {edges} edges ({gen_s:.1f}s)")
    try:
        queries = [(G, s[4], 0, None) for G, s in zip(Gs, specs)]
        for _ in range(args.warmup):
            ctx.count_batch(queries)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            eager_counts, _ = ctx.count_batch(queries)
        t_eager = (time.perf_counter() - t0) / args.steps
        t0 = time.perf_counter()
        tuned, tuned_ms = ctx.autotune_batch(queries, reps=2)
        t_tune = time.perf_counter() - t0
        log(f"autotuned {tuned} ({tuned_ms:.2f} ms, {t_tune:.1f} s)")
        with ctx.prepare_batch(queries) as B:
            for _ in range(args.warmup):
                B.run()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                batch_counts, batch_ms = B.run()
            t_batch = (time.perf_counter() - t0) / args.steps
        t0 = time.perf_counter()
        plain_counts = []
        plain_ms = {}
        for k, (G, p, _, _) in enumerate(queries):
            c, kms = G.count(p)
            plain_counts.append(c)
            plain_ms[p] = plain_ms.get(p, 0.0) + kms
            if (k + 1) % 256 == 0:
                log(f"plain {k + 1}/{len(queries)} ({time.perf_counter() - t0:.1f}s)")
        t_plain = time.perf_counter() - t0
        for name, got in (("batched", batch_counts), ("eager batched", eager_counts)):
            if got != plain_counts:
                bad = [i for i, (x, y) in enumerate(zip(got, plain_counts)) if x != y]
                raise AssertionError(f"{name} counts differ from plain counts at {bad[:10]}")
        oracle_checked = 0
        if args.batch_graphs == 2048 and args.batch_max_n == 262144:
            for ent in json.load(open(os.path.join(GOLDEN, "batch2048.json")))["sample"]:
                if plain_counts[ent["index"]] != ent["count"]:
                    raise AssertionError(f"batch graph {ent['index']}: {plain_counts[ent['index']]} != oracle {ent['count']}")
                oracle_checked += 1
    finally:
        for G in Gs:
            G.free()
        if own:
            ctx.close()
    nq = len(queries)
    knobs = {k: tuned[k] for k in ("batch_hub_deg", "batch_root_chunk", "batch_edge_chunk", "batch_blocks_per_cu")}
    return {"metric": METRIC, "value": round(nq / t_batch, 3), "unit": "queries/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_batch * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
            "config": {"workload": f"batched: {nq} graphs (half ER mean degree U[8,64], half BA m U{{4..16}}, "
                                   f"n log-uniform [1k,{args.batch_max_n}]), pattern uniform from "
                                   "{TC,P13,P1,P2}; autotuned knobs, all launches of a step replayed from one "
                                   "captured HIP graph",
                       "edges": edges, "autotuned": knobs, "autotune_s": round(t_tune, 2),
                       "eager_batched_queries_per_s": round(nq / t_eager, 3),
                       "plain_queries_per_s": round(nq / t_plain, 3),
                       "plain_kernel_ms_by_pattern": {p: round(v, 1) for p, v in sorted(plain_ms.items())},
                       "speedup_vs_plain": round(t_plain / t_batch, 3), "counts_equal_plain": True,
                       "counts_checked_against_oracle": oracle_checked, "host_build_and_upload_s": round(gen_s, 1)}}


# Largest R-MAT scale a served pattern draws its graph from (the request's
# graph is uniform among the resident graphs within the cap).  Patterns that
# no resident scale admits are not drawn; both are stated on the line.
# The caps keep one plain query within about a second (measured whole-graph
# times on one MI355X: P1 0.78 s at scale 20 and 9.2 s at 22; P2 1.27 s at 22
# and 9.4 s at 24; P3 0.39 s at 20; P5/P7/P9/P11/P12/P15 1.6-3.6 s at 18; P4
# 2.4 s at 16; P6, P8, P10 and P16 take seconds already at scales 12-16).
SERVE_CAP = {"P1": 20, "P2": 22, "P3": 20, "P13": 24, "P14": 24, "TC": 24,
             "P5": 18, "P7": 18, "P9": 18, "P11": 18, "P12": 18, "P15": 18,
             "P4": 0, "P6": 0, "P8": 0, "P10": 0, "P16": 0}
SERVE_PATTERNS = [f"P{i}" for i in range(1, 17)]


def serve_graph_scales(min_scale=18, max_scale=24, n=32):
    rng = np.random.default_rng(SEED)
    return [int(rng.integers(min_scale, max_scale + 1)) for _ in range(n)]


def serve_requests(scales, n_req, rate, seed):
    """Poisson arrivals at `rate`; each request a uniform pattern id of
    P1..P16 among those a resident scale admits, on a uniform graph within the
    pattern's cap."""
    rng = np.random.default_rng(seed)
    pats = [p for p in SERVE_PATTERNS if any(sc <= SERVE_CAP[p] for sc in scales)]
    reqs = []
    for _ in range(n_req):
        p = pats[int(rng.integers(0, len(pats)))]
        ok = [i for i, sc in enumerate(scales) if sc <= SERVE_CAP[p]]
        reqs.append((int(rng.choice(ok)), p))
    return reqs, np.cumsum(rng.exponential(1.0 / rate, n_req)), pats


def run_served(ctx, Gs, reqs, arrivals, lanes=0):
    res = [None] * len(reqs)
    with J.Server(ctx, heavy_lanes=lanes) as srv:
        tickets = [None] * len(reqs)
        start = time.perf_counter()

        def submitter():
            for k, (gi, p) in enumerate(reqs):
                dt = arrivals[k] - (time.perf_counter() - start)
                if dt > 0:
                    time.sleep(dt)
                tickets[k] = srv.submit(Gs[gi], p)
        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        for k in range(len(reqs)):
            res[k] = srv.wait(tickets[k])
        span = time.perf_counter() - start
    return [r[0] for r in res], np.array([r[1] for r in res]) / 1e3, span


def run_plain(Gs, reqs, arrivals):
    counts, lat = [], []
    start = time.perf_counter()
    for k, (gi, p) in enumerate(reqs):
        dt = arrivals[k] - (time.perf_counter() - start)
        if dt > 0:
            time.sleep(dt)
        c, _ = Gs[gi].count(p)
        counts.append(c)
        lat.append((time.perf_counter() - start - arrivals[k]) * 1e3)
    return counts, np.array(lat), time.perf_counter() - start


def pct(a, q):
    return round(float(np.percentile(a, q)), 3)


def serving(ctx, rates, n_req, min_scale=18, max_scale=24):
    """The serving configuration: 32 resident R-MAT graphs (ef 16), Poisson
    arrivals at each rate; served (continuous batching) and plain (FIFO, one
    query per launch) runs of the same request stream; counts must agree."""
    scales = serve_graph_scales(min_scale, max_scale)
    t0 = time.time()
    Gs = [ctx.rmat_graph(sc, EF, 0.57, 0.19, 0.19, seed=splitmix(SEED, i)) for i, sc in enumerate(scales)]
    log(f"serving: built 32 R-MAT graphs on the device, scales {sorted(scales)} ({time.time() - t0:.1f}s)")
    # warm every (graph, pattern) pair's lazily built arrays once, untimed
    rates_out = {}
    pats = None
    try:
        warm = set()
        for rate in rates:
            reqs, arr, pats = serve_requests(scales, n_req, rate, seed=int(rate))
            for gi, p in reqs:
                if (gi, p) not in warm:
                    Gs[gi].count(p, (0, 1))
                    warm.add((gi, p))
            log(f"serving rate {rate}: {len(reqs)} requests")
            sc, slat, sspan = run_served(ctx, Gs, reqs, arr)
            pc, plat, pspan = run_plain(Gs, reqs, arr)
            if sc != pc:
                bad = [i for i, (x, y) in enumerate(zip(sc, pc)) if x != y]
                raise AssertionError(f"served counts differ from plain counts at rate {rate}: {bad[:10]}")
            rates_out[str(rate)] = {
                "served_qps": round(len(reqs) / sspan, 2), "served_p50_ms": pct(slat, 50),
                "served_p99_ms": pct(slat, 99), "plain_qps": round(len(reqs) / pspan, 2),
                "plain_p50_ms": pct(plat, 50), "plain_p99_ms": pct(plat, 99),
                "served_over_plain": round(pspan / sspan, 3)}
            log(f"serving rate {rate}: {rates_out[str(rate)]}")
    finally:
        for G in Gs:
            G.free()
    best = max(rates_out.values(), key=lambda r: r["served_qps"])
    return {"workload": "32 resident R-MAT graphs (ef 16, 0.57/0.19/0.19/0.05), scales "
                        f"{min_scale}-{max_scale} uniform; Poisson arrivals; request = uniform pattern id of P1..P16 "
                        "on a uniform graph within the pattern's scale cap; continuous batching vs FIFO plain launches",
            "graph_scales": sorted(scales), "patterns_drawn": pats,
            "scale_caps": {p: SERVE_CAP[p] for p in SERVE_PATTERNS},
            "not_drawn": [p for p in SERVE_PATTERNS if p not in (pats or [])],
            "requests_per_rate": n_req, "rates": rates_out,
            "best_served_qps": best["served_qps"], "best_served_over_plain": best["served_over_plain"]}


def other_configs(ctx):
    """Triangle count on R-MAT 14 (latency, launch-bound) and the induced
    4-vertex motifs on the 3.8M-vertex Chung-Lu graph, each against the
    oracle's frozen counts."""
    out = {}
    gold = json.load(open(os.path.join(GOLDEN, "counts.json")))["rmat14_s1"]["counts"]
    G = ctx.rmat_graph(14, 16, 0.57, 0.19, 0.19, seed=1)
    for _ in range(5):
        G.count("TC")
    lat, kms = [], []
    for _ in range(100):
        t0 = time.perf_counter()
        c, ms = G.count("TC")
        lat.append((time.perf_counter() - t0) * 1e3)
        kms.append(ms)
        assert c == gold["TC"], (c, gold["TC"])
    G.free()
    out["tc_rmat14"] = {"workload": "TC on R-MAT scale 14 (ef 16, seed 1), whole graph, one range",
                        "count": gold["TC"], "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
                        "kernel_ms": {"p50": pct(kms, 50), "p99": pct(kms, 99)}}
    cl = json.load(open(os.path.join(GOLDEN, "chunglu.json")))
    t0 = time.time()
    g = J.chung_lu(cl["n"], cl["target_edges"], cl["gamma"], cl["max_weight"], seed=cl["seed"])
    G = ctx.upload(g)
    gen_s = time.time() - t0
    del g
    res = {}
    for p in ("I_C4", "I_DIAMOND", "I_TAILTRI", "I_STAR3", "P1", "P2", "P13"):
        G.count(p)  # builds the per-graph arrays once
        t0 = time.perf_counter()
        c, ms = G.count(p)
        wall = (time.perf_counter() - t0) * 1e3
        assert c == cl["counts"][p], (p, c, cl["counts"][p])
        res[p] = {"count": c, "kernel_ms": round(ms, 3), "wall_ms": round(wall, 3)}
    G.free()
    out["induced_chunglu"] = {"workload": f"vertex-induced 4-vertex motifs, Chung-Lu n={cl['nv']} "
                                          f"({cl['sym_nnz'] // 2} edges, gamma 2.3, max weight 800, seed 1)",
                              "host_build_s": round(gen_s, 2), "queries": res}
    return out


def profile_workload(kind):
    """One pass of the kernels a rocprofv3 summary in profiles/ covers (no
    timing loop): `sym`: P2, P3, P6, P9 on the 3.8M-vertex Chung-Lu graph and
    P2, P3, P9 on R-MAT 18 (P6 there costs minutes); `wide`: TC, P13, P14 on
    R-MAT 24, whose 8,382 roots of DAG out-degree 1025..1775 run in the
    2048-wide tier."""
    out = []
    with J.Context(0) as ctx:
        if kind == "sym":
            cl = json.load(open(os.path.join(GOLDEN, "chunglu.json")))
            g = J.chung_lu(cl["n"], cl["target_edges"], cl["gamma"], cl["max_weight"], seed=cl["seed"])
            graphs = [("chung-lu 3.8M", ctx.upload(g), ("P2", "P3", "P6", "P9")),
                      ("R-MAT 18", ctx.rmat_graph(18, EF, 0.57, 0.19, 0.19, seed=SEED), ("P2", "P3", "P9"))]
            del g
        else:
            graphs = [("R-MAT 24", ctx.rmat_graph(24, EF, 0.57, 0.19, 0.19, seed=SEED), ("TC", "P13", "P14"))]
        for name, G, pats in graphs:
            for p in pats:
                G.count(p, (0, 1))  # builds the per-graph arrays outside the profiled pass
                c, ms = G.count(p)
                out.append({"graph": name, "pattern": p, "count": c, "kernel_ms": round(ms, 3),
                            "kernels": ctx.last_profile()})
                log(f"{name} {p}: {c} ({ms:.2f} ms)")
            G.free()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-other", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--scale", type=int, default=SCALE)
    ap.add_argument("--profile", action="store_true",
                    help="only the counting kernels of one step (for rocprofv3 counter runs)")
    ap.add_argument("--workload", choices=["cliques", "batch", "serve", "profile-sym", "profile-wide"],
                    default="cliques")
    ap.add_argument("--batch-graphs", type=int, default=2048)
    ap.add_argument("--batch-max-n", type=int, default=262144)
    ap.add_argument("--serve-min-scale", type=int, default=18)
    ap.add_argument("--serve-max-scale", type=int, default=24)
    ap.add_argument("--rates", type=float, nargs="+", default=[1000])
    ap.add_argument("--requests", type=int, default=32)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="one plain 1 GiB device-to-device copy kernel (scale factor for FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": x.numel() * 4}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (a count shards by root range, "
              "each GPU adding its partial count). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        print(json.dumps(workload_batch(args)), flush=True)
        return
    if args.workload.startswith("profile-"):
        print(json.dumps(profile_workload(args.workload[len("profile-"):])), flush=True)
        return
    if args.workload == "serve":
        with J.Context(0) as ctx:
            print(json.dumps(serving(ctx, args.rates, args.requests, args.serve_min_scale, args.serve_max_scale)),
                  flush=True)
        return
    gold = json.load(open(os.path.join(GOLDEN, "rmat22.json")))
    with J.Context(0) as ctx:
        # the graph is sampled, symmetrised, deduplicated and oriented on the
        # GPU (identical to the host construction), untimed
        t0 = time.time()
        G = ctx.rmat_graph(args.scale, EF, 0.57, 0.19, 0.19, seed=SEED)
        ingest_s = time.time() - t0
        info = G.info()
        q = info["nv"] // 4
        log(f"R-MAT scale {args.scale} built on the device in {ingest_s:.2f}s: {info} on {ctx.info()}")
        if args.profile:
            for p in QUERIES:
                c, ms = G.count(p, (0, q))
                print(json.dumps({"pattern": p, "count": c, "kernel_ms": ms,
                                  "kernels": ctx.last_profile()}), flush=True)
            G.free()
            return
        for w in range(args.warmup):
            for p in QUERIES:
                c, ms = G.count(p, (0, q))
                log(f"warmup {w} {p} count={c} {ms:.2f} ms")
        counts = {}
        lat = {p: [] for p in QUERIES}
        kern = {}  # (pattern, kernel index, name) -> list of ms
        t_start = time.perf_counter()
        for s in range(args.steps):
            for p in QUERIES:
                ts = time.perf_counter()
                c, ms = G.count(p, (0, q))
                lat[p].append((time.perf_counter() - ts) * 1e3)
                if counts.setdefault(p, c) != c:
                    raise AssertionError(f"{p} count changed between steps: {counts[p]} != {c}")
                for i, (name, kms) in enumerate(ctx.last_profile()):
                    kern.setdefault((p, i, name), []).append(kms)
            log(f"step {s} done")
        elapsed = time.perf_counter() - t_start
        n_queries = args.steps * len(QUERIES)
        value = n_queries / elapsed
        # counts against the oracle's frozen counts (whole quarter for P13;
        # P14 over the quarter is checked by its root windows in the tests)
        checked = {}
        if args.scale == SCALE:
            for p, want in gold.get("quarter0", {}).items():
                if p in counts:
                    if counts[p] != want:
                        raise AssertionError(f"{p} count {counts[p]} differs from the oracle's {want}")
                    checked[p] = "oracle, whole quarter"
        # dominant tier: one launch, or the build and count launches of the
        # split 1024-wide tier; its time is the sum of their average times
        tiers = {}
        for (p, i, n), v in kern.items():
            tiers.setdefault((p, n[:-len("_build")] if n.endswith("_build") else n), []).append((n, float(np.mean(v))))
        (dp, dname), dk = max(tiers.items(), key=lambda kv: sum(ms for _, ms in kv[1]))
        avg_ms = sum(ms for _, ms in dk)
        dag = G.download(dag=True)
        G.free()
        t0 = time.time()
        per_tier = {p: design_bytes_by_tier(dag.rowptr, dag.col, 0, q, CLIQUE_K[p]) for p in QUERIES}
        log(f"design bytes per tier ({time.time() - t0:.1f}s): {per_tier}")
        dbytes = per_tier[dp][dname]["bytes"]
        achieved = dbytes / (avg_ms * 1e-3) / 1e9
        if achieved > HBM_PEAK_GBS:
            raise AssertionError(f"design bytes exceed the HBM peak: {achieved} GB/s")
        traffic = measured_traffic([n for n, _ in dk], dp)
        query_ms = {p: float(np.mean([sum(v[k] for (pp, _, _), v in kern.items() if pp == p)
                                      for k in range(args.steps)])) for p in QUERIES}
        query_bytes = {p: sum(t["bytes"] for t in per_tier[p].values()) for p in QUERIES}
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "kernel": f"{dname} ({dp})", "launches": {n: round(ms, 3) for n, ms in dk},
                    "kernel_avg_ms": round(avg_ms, 3), "algorithmic_bytes": dbytes,
                    "roots": per_tier[dp][dname]["roots"],
                    "per_query": {p: {"design_bytes": query_bytes[p], "device_ms": round(query_ms[p], 3),
                                      "GBps": round(query_bytes[p] / query_ms[p] / 1e6, 1),
                                      "frac": round(query_bytes[p] / query_ms[p] / 1e6 / HBM_PEAK_GBS, 4)}
                                  for p in QUERIES},
                    "traffic_source": "profiles/clique_p13_p14_rmat22_q0_traffic.json" if traffic else
                                      "not measured for these kernel sources"}
        line = {
            "metric": METRIC, "value": round(value, 4), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(elapsed * 1e3 / args.steps, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
            "config": {"workload": f"P13+P14 count, roots [0,2^{args.scale - 2}) of R-MAT scale {args.scale} "
                                   f"ef {EF} (0.57/0.19/0.19/0.05, seed {SEED}), degree DAG",
                       "nv": info["nv"], "sym_nnz": info["sym_nnz"], "dag_nnz": info["dag_nnz"],
                       "dag_max_deg": info["dag_max_deg"], "counts": counts, "counts_checked": checked,
                       "latency_ms": {p: {"p50": pct(v, 50), "p99": pct(v, 99)} for p, v in lat.items()},
                       "kernels_ms": {f"{p}:{n}": round(float(np.mean(v)), 3) for (p, i, n), v in kern.items()},
                       "device_ingest_s": round(ingest_s, 3)},
            "roofline": roofline,
        }
        if not args.no_other:
            log("other configurations ...")
            line["other_configs"] = other_configs(ctx)
        if not args.no_batch:
            log("batched workload ...")
            b = workload_batch(argparse.Namespace(**{**vars(args), "steps": 3, "warmup": 1}), ctx)
            line["batched"] = {"queries_per_s": b["value"], "ms_per_step": b["ms_per_step"], **b["config"]}
        if not args.no_serve:
            log("serving ...")
            line["serving"] = serving(ctx, args.rates, args.requests, args.serve_min_scale, args.serve_max_scale)
    if not args.no_cpu:
        log("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(dag, q)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
