"""Benchmark: 4-clique (P13) and 5-clique (P14) counting over one quarter of
the roots of a seeded R-MAT graph (scale 22, edge factor 16, a/b/c/d =
0.57/0.19/0.19/0.05, degree-ordered DAG, int32 CSR, uint64 counts).

A step is one P13 query plus one P14 query over roots [0, 2^20), each a
plain per-call launch that returns its exact count to the host.  `value` is
queries/s over the timed steps (inputs resident in HBM before timing).

Also reported on the same JSON line:
  roofline      the dominant tier (largest device time over the timed
                region, from HIP events on the library's stream; one launch,
                or the build and count launches of the split 1024-wide tier):
                work-model bytes of the tier (every neighbour-list fetch of
                the reference's loop nest costs 16 + 4 d(v) bytes) / its
                average time per query, against the 8 TB/s HBM peak;
                `traffic` is its measured HBM bytes (rocprofv3 FETCH_SIZE +
                WRITE_SIZE, profiles/);
  cpu_baseline  the CPU oracle (clean-room OpenMP restatement of the
                reference nests) timed on a seeded random sample of the same
                roots, scaled to the full query;
  latency       p50 / p99 per query kind over the timed steps.
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import jupiter_amd as J  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "queries/s served vs plain per-call kernel (bit-exact counts); p50/p99; GB/s"
SCALE, EF, SEED = 22, 16, 1
QUERIES = ("P13", "P14")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(dag, roots_end, frac_den=1024, seed=7):
    """Oracle on a seeded uniform sample of 1/frac_den of the roots, both
    query kinds; scaled by frac_den to the full query."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    rng = np.random.default_rng(seed)
    roots = np.sort(rng.choice(roots_end, size=roots_end // frac_den, replace=False)).astype(np.int32)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    t_total = 0.0
    for p in QUERIES:
        t0 = time.perf_counter()
        O.count_roots(dag.rowptr, dag.col, J.PATTERNS[p], roots, mode=1, threads=threads)
        t_total += time.perf_counter() - t0
    est = t_total * frac_den  # seconds for one P13 + one P14 over all roots
    return {"value": round(len(QUERIES) / est, 6), "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"{len(roots)} roots (uniform seeded 1/{frac_den} of [0,{roots_end})), P13+P14, "
                      f"{t_total:.1f}s measured, scaled x{frac_den}"}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "clique_p13_p14_rmat22_q0_traffic.json")


def measured_traffic(kernels, scale):
    """HBM bytes (read + written) of one pass of the launches `kernels` on
    the default workload, from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    summarised in profiles/ (scaled by the plain-copy calibration); None when
    one of them is not measured."""
    if scale != SCALE or not os.path.exists(TRAFFIC_FILE):
        return None
    tbl = json.load(open(TRAFFIC_FILE)).get("kernels", {})
    if any(k not in tbl for k in kernels):
        return None
    return int(sum(tbl[k]["read_bytes"] + tbl[k]["write_bytes"] for k in kernels))


def splitmix(seed, i):
    x = (seed * 0x9E3779B97F4A7C15 + i * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    x ^= x >> 31
    return (x * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF


def workload_batch(args):
    """Many independent graphs, each paired with one pattern from
    {TC, P13, P1, P2}: the batch autotuned, captured once in a HIP graph and
    replayed, vs one plain launch per query.  The default is the full
    configuration (2,048 graphs, n up to 262,144)."""
    rng = np.random.default_rng(SEED)
    specs = []
    t0 = time.time()
    edges = 0
    ctx = J.Context(0)
    Gs = []
    for i in range(args.batch_graphs):
        n = int(np.exp(rng.uniform(np.log(1000), np.log(args.batch_max_n))))
        sd = splitmix(SEED, i)
        if i % 2 == 0:
            md = float(rng.uniform(8, 64))
            c = J.erdos_renyi(n, md, seed=sd)
            kind = f"ER(n={n},deg={md:.1f})"
        else:
            m = int(rng.integers(4, 17))
            c = J.barabasi_albert(n, m, seed=sd)
            kind = f"BA(n={n},m={m})"
        specs.append((kind, ["TC", "P13", "P1", "P2"][int(rng.integers(0, 4))]))
        edges += c.nnz // 2
        Gs.append(ctx.upload(c))  # resident; the host copy is dropped
        if (i + 1) % 128 == 0:
            log(f"generated and uploaded {i + 1} graphs, {edges} edges ({time.time() - t0:.1f}s)")
    log(f"generated {len(Gs)} graphs, {edges} edges ({time.time() - t0:.1f}s)")
    with ctx:
        queries = [(G, p, 0, None) for G, (_, p) in zip(Gs, specs)]
        for _ in range(args.warmup):
            ctx.count_batch(queries)
        # eager batched launch with the default knobs
        t0 = time.perf_counter()
        for _ in range(args.steps):
            eager_counts, _ = ctx.count_batch(queries)
        t_eager = (time.perf_counter() - t0) / args.steps
        # autotune the knobs on this batch (not timed), then capture the batch
        # in a HIP graph and time its replays
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
        for G in Gs:
            G.free()
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
                       "speedup_vs_plain": round(t_plain / t_batch, 3), "counts_equal_plain": True}}


SERVE_CAP = {"TC": 99, "P2": 99, "P13": 99, "P3": 99, "P5": 99, "P7": 99, "P9": 99, "P14": 99,
             "P1": 14, "P11": 14, "P12": 14, "P4": 12, "P8": 12, "P15": 12, "P6": 11, "P10": 10, "P16": 10}


def workload_serve(args):
    """Poisson request stream over resident R-MAT graphs, each request a
    uniform pattern id (TC, P1..P16) on a uniform graph (within the
    pattern's scale cap), served with continuous batching vs sequential plain
    launches (FIFO, one query per launch).  Counts must be identical."""
    import threading
    rng = np.random.default_rng(SEED)
    scales = [int(rng.integers(args.serve_min_scale, args.serve_max_scale + 1)) for _ in range(32)]
    pats = ["TC"] + [f"P{i}" for i in range(1, 17)]
    out = {"rates": {}}
    with J.Context(0) as ctx:
        t0 = time.time()
        Gs = [ctx.rmat_graph(sc, EF, 0.57, 0.19, 0.19, seed=splitmix(SEED, i)) for i, sc in enumerate(scales)]
        log(f"built 32 R-MAT graphs on the device, scales {sorted(scales)} ({time.time() - t0:.1f}s)")
        for rate in args.rates:
            reqs = []
            for _ in range(args.requests):
                p = pats[int(rng.integers(0, len(pats)))]
                ok = [i for i, sc in enumerate(scales) if sc <= SERVE_CAP[p]]
                reqs.append((int(rng.choice(ok)), p))
            arrivals = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
            # served
            res = [None] * len(reqs)
            with J.Server(ctx) as srv:
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
            lat = np.array([r[1] for r in res]) / 1e3
            # sequential plain launches, FIFO from the same arrival times
            plain = []
            plat = []
            start = time.perf_counter()
            for k, (gi, p) in enumerate(reqs):
                dt = arrivals[k] - (time.perf_counter() - start)
                if dt > 0:
                    time.sleep(dt)
                c, _ = Gs[gi].count(p)
                plain.append(c)
                plat.append((time.perf_counter() - start - arrivals[k]) * 1e3)
            pspan = time.perf_counter() - start
            if [r[0] for r in res] != plain:
                raise AssertionError(f"served counts differ from plain counts at rate {rate}")
            out["rates"][str(rate)] = {
                "served_qps": round(len(reqs) / span, 2), "served_p50_ms": round(float(np.percentile(lat, 50)), 3),
                "served_p99_ms": round(float(np.percentile(lat, 99)), 3),
                "plain_qps": round(len(reqs) / pspan, 2),
                "plain_p50_ms": round(float(np.percentile(plat, 50)), 3),
                "plain_p99_ms": round(float(np.percentile(plat, 99)), 3)}
            log(f"rate {rate}: {out['rates'][str(rate)]}")
        for G in Gs:
            G.free()
    best = max(out["rates"].values(), key=lambda r: r["served_qps"])
    return {"metric": METRIC, "value": best["served_qps"], "unit": "queries/s", "n_gpus": 1, "steps": len(args.rates),
            "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "int32", "data": "synthetic",
            "config": {"workload": f"serving: 32 resident R-MAT graphs scales {args.serve_min_scale}-"
                                   f"{args.serve_max_scale} (ef 16), Poisson arrivals, uniform pattern id "
                                   "(TC,P1..P16) with per-pattern scale caps, continuous batching vs "
                                   "sequential plain launches", "scale_caps": SERVE_CAP,
                       "requests_per_rate": args.requests, **out}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--scale", type=int, default=SCALE)
    ap.add_argument("--profile", action="store_true",
                    help="only the counting kernels of one step (for rocprofv3 counter runs)")
    ap.add_argument("--workload", choices=["cliques", "batch", "serve"], default="cliques")
    ap.add_argument("--batch-graphs", type=int, default=2048)
    ap.add_argument("--batch-max-n", type=int, default=262144)
    ap.add_argument("--serve-min-scale", type=int, default=10)
    ap.add_argument("--serve-max-scale", type=int, default=16)
    ap.add_argument("--rates", type=float, nargs="+", default=[100, 500, 2000])
    ap.add_argument("--requests", type=int, default=200)
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
    if args.workload == "serve":
        print(json.dumps(workload_serve(args)), flush=True)
        return
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
        # dominant tier and its work-model bytes.  A tier is one launch, or
        # two for the split 1024-wide tier (its "_build" launch writes the
        # rows that the count launch reads): its time is the sum of its
        # launches' average times
        tiers = {}
        for (p, i, n), v in kern.items():
            tiers.setdefault((p, n[:-len("_build")] if n.endswith("_build") else n), []).append((n, float(np.mean(v))))
        (dp, dname), dk = max(tiers.items(), key=lambda kv: sum(ms for _, ms in kv[1]))
        avg_ms = sum(ms for _, ms in dk)
        G.count(dp, (0, q))
        names = [n for n, _ in ctx.last_profile() if not n.endswith("_build")]
        # the work model runs the fused per-tier kernels: match by tier name
        by_name = dict(zip(names, G.work_bytes_split(dp, (0, q))))
        dbytes = by_name[dname]
        achieved = dbytes / (avg_ms * 1e-3) / 1e9
        total_bytes = {p: G.work_bytes(p, (0, q)) for p in QUERIES}
        dag = None if args.no_cpu else G.download(dag=True)
        G.free()
    kname = f"{dname} ({dp})"
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic([f"{n} ({dp})" for n, _ in dk], args.scale),
                "kernel": kname, "launches": {n: round(ms, 3) for n, ms in dk}, "kernel_avg_ms": round(avg_ms, 3),
                "algorithmic_bytes": dbytes}
    line = {
        "metric": METRIC, "value": round(value, 4), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed * 1e3 / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"P13+P14 count, roots [0,2^{args.scale - 2}) of R-MAT scale {args.scale} "
                               f"ef {EF} (0.57/0.19/0.19/0.05, seed {SEED}), degree DAG",
                   "nv": info["nv"], "sym_nnz": info["sym_nnz"], "dag_nnz": info["dag_nnz"],
                   "dag_max_deg": info["dag_max_deg"], "counts": counts,
                   "latency_ms": {p: {"p50": round(float(np.percentile(v, 50)), 3),
                                      "p99": round(float(np.percentile(v, 99)), 3)} for p, v in lat.items()},
                   "kernels_ms": {f"{p}:{n}": round(float(np.mean(v)), 3) for (p, i, n), v in kern.items()},
                   "work_bytes": total_bytes, "device_ingest_s": round(ingest_s, 3)},
        "roofline": roofline,
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(dag, q)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
