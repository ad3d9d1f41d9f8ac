#!/usr/bin/env python3
"""Closure throughput on one MI355X.

Workload: the in-core pointer-alias closure -- 1,000,000 vertices from the
seeded alias generator (graspan_g.gen.alias, seed 0: procedures of U[20,200]
vertices, Poisson(1.5) assign edges, 10% deref edges, call sites to Zipf(1.1)
callees inside modules of 8 procedures, reverse A/D edges; ~3.2M input edges)
under the alias grammar (L ::= eps | L A | L M; R ::= eps | R a | R M;
V ::= L R; T ::= D V; M ::= T d), filter 0.

One step = one full fixpoint of that graph with its input already resident in
HBM. value = new closure edges per second = (|closure| - |dedup input|) / t,
summed over the timed steps.

roofline: the join/dedup/difference kernel (k_expand), timed live with HIP
events on the engine's stream. Its algorithmic bytes per round are
4*P + 4*O + 4*N (P probe reads: key elements scanned plus every candidate-list
element read; O old/delta row elements read for the difference; N new edges
written), from the engine's counters, which the parity tests hold equal to the
oracle's.

cpu_baseline: the C++/OpenMP oracle on a bounded sample of the same workload
(the same generator at 100,000 vertices, run to its fixpoint), new closure
edges per second on the host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "graspan-g_amd"))

HBM_PEAK_GBS = 8000.0
WORKLOAD_N = 1_000_000
SAMPLE_N = 100_000


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def traffic_from_profiles():
    """HBM bytes per k_expand launch measured by rocprofv3 PMC counters, if a
    summary for this workload is committed."""
    p = os.path.join(ROOT, "profiles", "expand_alias1m_traffic.json")
    try:
        with open(p) as f:
            return json.load(f).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # checker / baseline only
    from graspan_g import gen

    g = gen.alias(SAMPLE_N, seed=0)
    threads = os.cpu_count() or 1
    r = oracle.closure_of(g, want_triples=False, threads=threads)
    new = r.total - r.input_unique
    return {"value": new / r.seconds, "unit": "closure edges/s", "cores": threads, "kind": "port",
            "sample": f"alias generator n={SAMPLE_N} seed 0 run to its fixpoint ({r.total} closure edges, "
                      f"{r.seconds:.2f} s, OpenMP threads={threads})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=WORKLOAD_N, help="vertex count (default: the 1M-vertex workload)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--verbose", action="store_true", help="per-round engine statistics on stderr")
    ap.add_argument("--mode", choices=("single", "batch", "serve", "ooc", "dataflow"), default="single",
                    help="single: the 1M-vertex alias closure (default); batch: graphs/s of the batched closure "
                         "against the same engine one graph at a time; serve: open-loop Poisson arrivals through "
                         "the serving loop, p50/p99 latency and graphs/s; ooc: out-of-core closure of the "
                         "40M-vertex block dataflow graph under a 64 GB device budget; dataflow: the 10,000-vertex "
                         "dataflow closure")
    ap.add_argument("--batch-stride", type=int, default=16,
                    help="batch mode: close every k-th graph of the 4096-graph set")
    ap.add_argument("--serve-requests", type=int, default=120)
    ap.add_argument("--serve-rate", type=float, default=4.0, help="serve mode: arrivals per second (Poisson)")
    ap.add_argument("--serve-min-n", type=int, default=1_000)
    ap.add_argument("--serve-max-n", type=int, default=2_000_000)
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--ooc-n", type=int, default=40_000_000)
    ap.add_argument("--ooc-partitions", type=int, default=64)
    ap.add_argument("--ooc-budget-gb", type=float, default=64.0)
    ap.add_argument("--ooc-check", action="store_true",
                    help="ooc mode: also close the graph in core (no budget) and compare checksums")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (calibrates rocprofv3 FETCH_SIZE / "
                         "WRITE_SIZE against known bytes) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        nbytes = args.copy_calibration << 20
        src = torch.empty(nbytes // 4, dtype=torch.int32, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": nbytes}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import graspan_g as gg
    from graspan_g import gen

    if args.mode == "batch":
        return bench_batch(args)
    if args.mode == "serve":
        return bench_serve(args)
    if args.mode == "ooc":
        return bench_ooc(args)
    if args.mode == "dataflow":
        return bench_dataflow(args)

    t = time.time()
    g = gen.alias(args.n, seed=0)
    log(f"generated alias n={g.n} m={g.m} in {time.time() - t:.1f}s")
    ctx = gg.Context(verbose=args.verbose)
    ctx.set_spec(g)
    for i in range(args.warmup):
        st = ctx.run(0)
        log(f"warmup {i}: {st.fixpoint_ms:.1f} ms, closure {st.total_edges}, rounds {st.rounds}")
    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = ctx.run(0)
        stats.append(st.as_dict())
        log(f"step {i}: {st.fixpoint_ms:.1f} ms (expand {st.expand_ms:.1f}, merge {st.merge_ms:.1f}, "
            f"compact {st.compact_ms:.1f}, other {st.other_ms:.1f}) closure {st.total_edges}")
    wall = time.perf_counter() - t0
    new_edges = sum(s["total_edges"] - s["input_unique"] for s in stats)
    value = new_edges / wall
    s0 = stats[-1]
    exp_ms = sum(s["expand_ms"] for s in stats)
    launches = sum(s["expand_launches"] for s in stats)
    alg = sum(4 * s["probe_reads"] + 4 * s["old_reads"] + 4 * s["new_edges"] for s in stats)
    achieved = alg / (exp_ms / 1e3) / 1e9 if exp_ms > 0 else 0.0
    traffic = traffic_from_profiles()
    line = {
        "metric": "closure edges/sec",
        "value": value,
        "unit": "closure edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"alias closure, n={g.n}, {g.m} input edges, alias grammar, filter 0",
                   "closure_edges": s0["total_edges"], "input_unique": s0["input_unique"],
                   "rounds": s0["rounds"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_expand (join+dedup+difference)",
                     "alg_bytes_per_launch": alg / max(1, launches),
                     "avg_launch_ms": exp_ms / max(1, launches)},
        "breakdown_ms": {k: sum(s[k] for s in stats) / len(stats)
                         for k in ("fixpoint_ms", "expand_ms", "compact_ms", "merge_ms", "other_ms")},
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on the n=100000 sample)")
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


def bench_batch(args):
    """graphs/s of the batched closure (every graph of the sample in one
    launch group, one launch sequence per round) against the same kernels run
    one graph at a time, on a uniform sample of the 4096-graph set (vertex
    counts log-uniform in [500, 200000], alternating dataflow / alias)."""
    import graspan_g as gg
    from graspan_g import gen

    t = time.time()
    sizes = gen.log_uniform_sizes(4096, 500, 200_000, seed=0)
    idx = list(range(0, 4096, args.batch_stride))
    graphs = [gen.alias_as_dataflow(int(sizes[i]), seed=i) if i % 2 == 0 else gen.alias(int(sizes[i]), seed=i)
              for i in idx]
    nv = sum(g.n for g in graphs)
    log(f"generated {len(graphs)} graphs ({nv} vertices) in {time.time() - t:.1f}s")
    ctx = gg.Context()
    b = gg.Batch(ctx)
    for g in graphs:
        b.add_spec(g)
    b.run()  # warmup
    best_b = None
    for i in range(max(1, args.steps)):
        t0 = time.perf_counter()
        st = b.run()
        dt = time.perf_counter() - t0
        best_b = dt if best_b is None else min(best_b, dt)
        log(f"batched step {i}: {dt * 1e3:.1f} ms, {st.rounds} rounds, closure {st.total_edges}")
    batched = [b.result(i) for i in range(len(graphs))]
    one = gg.Context()
    t0 = time.perf_counter()
    single = []
    for g in graphs:
        one.set_spec(g)
        one.run(0)
        single.append((one.counts()[1], one.checksum()))
    dt_s = time.perf_counter() - t0
    log(f"one graph at a time: {dt_s * 1e3:.1f} ms")
    for i, (tot, cs) in enumerate(single):
        if (tot, cs) != (batched[i][1], batched[i][2]):
            raise SystemExit(f"graph {idx[i]}: batched result differs from the single-graph result")
    line = {
        "metric": "graphs/sec (batched closure)",
        "value": len(graphs) / best_b,
        "unit": "graphs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": best_b * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"batched closure of graphs {idx[0]}, {idx[1]}, ... (every {args.batch_stride}th) "
                               f"of the 4096-graph set: {len(graphs)} graphs, {nv} vertices",
                   "one_at_a_time_graphs_per_s": len(graphs) / dt_s,
                   "speedup_vs_one_at_a_time": dt_s / best_b,
                   "closure_edges": int(sum(x[1] for x in batched))},
    }
    print(json.dumps(line), flush=True)


def bench_serve(args):
    """Serving loop under open-loop Poisson arrivals. Requests are graphs of
    the alias generator with log-uniform vertex counts, grammar 50/50
    dataflow / alias, admitted by estimated footprint under the budget. A
    seeded 1% sample (at least two requests) is checked against the oracle."""
    import numpy as np

    import graspan_g as gg
    from graspan_g import gen

    N = args.serve_requests
    rng = np.random.default_rng(0)
    sizes = np.exp(rng.uniform(np.log(args.serve_min_n), np.log(args.serve_max_n), size=N)).astype(np.int64)
    use_alias = rng.random(N) < 0.5
    gaps = rng.exponential(1.0 / args.serve_rate, size=N)
    arrive = np.cumsum(gaps) - gaps[0]
    t = time.time()
    graphs = [gen.alias(int(nv), seed=1000 + i) if use_alias[i] else gen.alias_as_dataflow(int(nv), seed=1000 + i)
              for i, nv in enumerate(sizes)]
    log(f"generated {N} requests ({int(sizes.sum())} vertices) in {time.time() - t:.1f}s")
    srv = gg.Server(slots=args.serve_slots, budget_bytes=int(args.serve_budget_gb * (1 << 30)))
    # warm the device up with one small request outside the measurement
    w = srv.submit_spec(gen.alias(2000, seed=1))
    srv.wait(w)
    t0 = srv.now_ms()
    ids = []
    for i, g in enumerate(graphs):
        delay = (t0 + arrive[i] * 1e3 - srv.now_ms()) / 1e3
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit_spec(g))
    res = []
    for i, rid in enumerate(ids):
        r, per = srv.wait(rid)
        res.append((r, per))
    srv.close()
    lat = np.array([r.done_ms - r.submit_ms for r, _ in res])
    span = max(r.done_ms for r, _ in res) - min(r.submit_ms for r, _ in res)
    edges = sum(r.total_edges for r, _ in res)
    log(f"served {N} requests in {span:.0f} ms; p50 {np.percentile(lat, 50):.0f} ms, p99 {np.percentile(lat, 99):.0f} ms")
    # seeded 1% sample against the oracle
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # checker only
    sample = sorted(np.random.default_rng(1).choice(N, size=max(2, N // 100), replace=False).tolist())
    for i in sample:
        ref = oracle.closure_of(graphs[i], want_triples=False)
        r, per = res[i]
        if (r.total_edges, r.checksum) != (ref.total, ref.checksum) or not np.array_equal(per, ref.counts):
            raise SystemExit(f"request {i}: served result differs from the oracle")
    log(f"oracle check of requests {sample}: equal")
    line = {
        "metric": "graphs/sec (serving loop)",
        "value": N / (span / 1e3),
        "unit": "graphs/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": span,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"{N} requests, Poisson {args.serve_rate}/s, n log-uniform "
                               f"[{args.serve_min_n}, {args.serve_max_n}], grammar 50/50, "
                               f"{args.serve_slots} slots, {args.serve_budget_gb} GB budget",
                   "p50_latency_ms": float(np.percentile(lat, 50)),
                   "p99_latency_ms": float(np.percentile(lat, 99)),
                   "mean_batch_size": float(np.mean([r.batch_size for r, _ in res])),
                   "closure_edges": int(edges),
                   "closure_edges_per_s": edges / (span / 1e3),
                   "oracle_checked_requests": sample},
    }
    print(json.dumps(line), flush=True)


def bench_ooc(args):
    """Out-of-core closure: dataflow grammar over gen.dataflow_blocks (the
    40,000,000-vertex workload, ~1.2*10^8 e edges, ~7.8*10^9 closure edges),
    64 contiguous partitions, device budget 64 GB; partitions outside the
    current pair live in pinned host DRAM."""
    import graspan_g as gg
    from graspan_g import gen

    t = time.time()
    g = gen.dataflow_blocks(args.ooc_n, seed=0)
    log(f"generated blocks n={g.n} m={g.m} in {time.time() - t:.1f}s")
    ctx = gg.Context(device_budget=int(args.ooc_budget_gb * (1 << 30)), verbose=True)  # progress on stderr
    ctx.set_spec(g)
    t0 = time.perf_counter()
    st = ctx.run_out_of_core(args.ooc_partitions)
    wall = time.perf_counter() - t0
    per, total = ctx.counts()
    cs = ctx.checksum()
    log(f"out-of-core: {wall * 1e3:.0f} ms, closure {total}, pairs run {st.pairs_run} skipped {st.pairs_skipped}, "
        f"spilled {st.spill_bytes / 1e9:.1f} GB, device peak {st.device_bytes_peak / 1e9:.1f} GB")
    check = None
    if args.ooc_check:
        ctx.release()
        inc = gg.Context()
        inc.set_spec(g)
        s2 = inc.run(0)
        check = {"in_core_total": int(s2.total_edges), "in_core_checksum_equal": inc.checksum() == cs,
                 "in_core_ms": s2.fixpoint_ms}
        if not check["in_core_checksum_equal"] or s2.total_edges != total:
            raise SystemExit("out-of-core closure differs from the in-core closure")
        log(f"in-core check: equal ({s2.fixpoint_ms:.0f} ms in core)")
    line = {
        "metric": "closure edges/sec (out-of-core)",
        "value": (total - st.input_unique) / wall,
        "unit": "closure edges/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"dataflow blocks n={g.n}, {g.m} e edges, {args.ooc_partitions} partitions, "
                               f"{args.ooc_budget_gb} GB device budget",
                   "closure_edges": int(total), "pairs_run": st.pairs_run, "pairs_skipped": st.pairs_skipped,
                   "spill_bytes": int(st.spill_bytes), "device_bytes_peak": int(st.device_bytes_peak),
                   "supersteps": st.supersteps, "check": check},
    }
    print(json.dumps(line), flush=True)


def bench_dataflow(args):
    """The 10,000-vertex plain dataflow closure (grammar N ::= e | N e,
    filter 0): closure edges/s, input resident in HBM."""
    import graspan_g as gg
    from graspan_g import gen

    g = gen.dataflow(10_000, seed=0)
    ctx = gg.Context(verbose=args.verbose)
    ctx.set_spec(g)
    for _ in range(args.warmup):
        ctx.run(0)
    t0 = time.perf_counter()
    stats = [ctx.run(0) for _ in range(args.steps)]
    wall = time.perf_counter() - t0
    new = sum(s.total_edges - s.input_unique for s in stats)
    line = {"metric": "closure edges/sec", "value": new / wall, "unit": "closure edges/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
            "config": {"workload": f"dataflow closure n={g.n}, {g.m} e edges, filter 0",
                       "closure_edges": int(stats[-1].total_edges), "rounds": stats[-1].rounds},
            "breakdown_ms": {k: getattr(stats[-1], k) for k in
                             ("fixpoint_ms", "expand_ms", "compact_ms", "merge_ms", "other_ms")}}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
