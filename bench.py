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
summed over the timed steps. The engine evaluates the least-fixpoint rule plan
(gg_grammar_plan_fixpoint: V ::= L R as V ::= R | A V | M V, same closure);
rule_table_as_given (extra key) times the same closure with the rule table as
given (GG_STAR_REWRITE=0).

roofline: the expand stage -- the join/dedup/difference kernels (k_dedup_light,
the first round's dedup of the input, one lane per row; k_expand_light,
one lane per row, in rounds with few candidates per row; k_expand_group, four
rows per wave in one guessed window, in the other rounds; k_expand, one wave
per row, for the rows either of them leaves) -- timed live with HIP events on the
engine's stream around each launch. Its VALU instructions and HBM bytes per
launch come from the committed rocprofv3 counter summary of the same workload
(profiles/expand_alias1m_counters.*, made by summarize_counters.py). The bound is the resource it uses the
larger share of: VALU issue (instructions x 64 lanes over the live launch
time, against 39.3 T lane-ops/s) -- the kernel is bound there -- or HBM
(bytes over the launch time, against 8 TB/s); both shares are reported. The
algorithmic byte models ride beside them: 4*P + 4*O + 4*N for this fused
kernel (P probe reads: key elements scanned plus every candidate-list element
read; O old/delta row elements read for the difference; N new edges written)
and 4*P + 8*K + 4*O + 8*N for a design that materialises its K candidates;
P, K, O, N are the engine's counters, which the parity tests hold equal to
the oracle's.

batch (extra key): the 4096-graph batched set, closed in launch groups,
against the same kernels one graph at a time (graphs/s and their ratio).
serve (extra key): the serving loop under open-loop Poisson arrivals at 200
requests/s (p50/p99 latency, graphs/s), with a 1% oracle sample.
ooc (extra key): the out-of-core closure of the 40,000,000-vertex block
dataflow graph (64 partitions under a 64 GB device budget, partitions spilled
to pinned host DRAM), closure edges/s with the input starting on the host.

cpu_baseline: the C++/OpenMP oracle on the same 1,000,000-vertex graph, run to
its fixpoint on the host's cores (OMP_NUM_THREADS of them), new closure edges
per second.
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
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12  # CUs x SIMDs x lanes x peak clock
WORKLOAD_N = 1_000_000


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def counters_from_profiles():
    """Per-launch hardware counters of the expand stage on this workload, from the
    committed rocprofv3 summary (profiles/expand_alias1m_counters.json), or
    None if it is absent."""
    p = os.path.join(ROOT, "profiles", "expand_alias1m_counters.json")
    try:
        with open(p) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """Threads the host may use: OMP_NUM_THREADS where it is set (the CPU
    share of this process), else every CPU."""
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "")))
    except ValueError:
        return os.cpu_count() or 1


def cpu_baseline(g):
    """The C++/OpenMP oracle on the same generated graph, run to its fixpoint
    on the host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # checker / baseline only

    threads = host_threads()
    r = oracle.closure_of(g, want_triples=False, threads=threads)
    new = r.total - r.input_unique
    return {"value": new / r.seconds, "unit": "closure edges/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"the whole workload: {g.name} ({r.total} closure edges, {r.rounds.shape[0]} rounds) "
                      f"run to its fixpoint in {r.seconds:.2f} s with {threads} OpenMP threads"}


def roofline(stats):
    """Roofline object for the expand stage (k_dedup_light, k_expand_light,
    k_expand_group and k_expand). Its
    time is measured live (HIP events on the engine's stream around every
    launch); its HBM traffic and VALU work
    per launch come from the committed rocprofv3 counter summary of the same
    workload. Both byte models are reported: the fused kernel's own
    (4P + 4O + 4N: probe reads, old/delta reads for the difference, new edges
    written) and the 4P + 8K + 4O + 8N of a design that writes and re-reads
    every candidate (K) and every new edge."""
    exp_ms = sum(s["expand_ms"] for s in stats)
    launches = sum(s["expand_launches"] for s in stats)
    avg_s = exp_ms / 1e3 / max(1, launches)
    fused = sum(4 * s["probe_reads"] + 4 * s["old_reads"] + 4 * s["new_edges"] for s in stats) / max(1, launches)
    survey = sum(4 * s["probe_reads"] + 8 * s["candidates"] + 4 * s["old_reads"] + 8 * s["new_edges"]
                 for s in stats) / max(1, launches)
    out = {"kernel": "expand stage: k_dedup_light + k_expand_light + k_expand_group + k_expand (join+dedup+difference)",
           "avg_launch_ms": avg_s * 1e3,
           "launches_per_step": launches / max(1, len(stats)),
           "alg_bytes_per_launch_fused": fused, "alg_bytes_per_launch_4P8K4O8N": survey,
           "alg_gbs_fused": fused / avg_s / 1e9, "alg_gbs_4P8K4O8N": survey / avg_s / 1e9}
    c = counters_from_profiles()
    if not c:
        # no counter summary: the fused byte model against HBM peak
        out.update({"bound": "hbm", "achieved": out["alg_gbs_fused"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": out["alg_gbs_fused"] / HBM_PEAK_GBS, "traffic": None})
        return out
    traffic = c["hbm_bytes_per_launch"]
    hbm_gbs = traffic / avg_s / 1e9
    # VALU work: wave64 instructions x 64 lanes per second, against 256 CUs x
    # 4 SIMD16 at the 2.4 GHz peak engine clock
    valu = c["valu_insts_per_launch"] * 64 / avg_s / 1e12
    out.update({"hbm_achieved_gbs": hbm_gbs, "hbm_peak_gbs": HBM_PEAK_GBS, "hbm_frac": hbm_gbs / HBM_PEAK_GBS,
                "valu_achieved_tops": valu, "valu_peak_tops": VALU_PEAK_TOPS, "valu_frac": valu / VALU_PEAK_TOPS,
                "traffic": traffic, "counters_source": c.get("source")})
    # the bound is whichever resource the kernel uses the larger share of
    if out["valu_frac"] >= out["hbm_frac"]:
        out.update({"bound": "valu", "achieved": valu, "peak": VALU_PEAK_TOPS, "unit": "Tops/s (int32 VALU lane-ops)",
                    "frac": out["valu_frac"]})
    else:
        out.update({"bound": "hbm", "achieved": hbm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": out["hbm_frac"]})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=WORKLOAD_N, help="vertex count (default: the 1M-vertex workload)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-closure leg of the default line")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg of the default line")
    ap.add_argument("--no-ooc", action="store_true", help="skip the out-of-core leg of the default line")
    ap.add_argument("--no-as-given", action="store_true",
                    help="skip the rule-table-as-given comparison of the default line")
    ap.add_argument("--verbose", action="store_true", help="per-round engine statistics on stderr")
    ap.add_argument("--mode", choices=("single", "batch", "serve", "ooc", "dataflow"), default="single",
                    help="single: the 1M-vertex alias closure, with the batched and serving legs as extra keys "
                         "(default); batch: only the batched closure of the 4096-graph set against the same "
                         "kernels one graph at a time; serve: only the serving loop under open-loop Poisson "
                         "arrivals; ooc: out-of-core closure of the 40M-vertex block dataflow graph under a 64 GB "
                         "device budget; dataflow: the 10,000-vertex dataflow closure")
    ap.add_argument("--batch-stride", type=int, default=1,
                    help="batch leg: close every k-th graph of the 4096-graph set (1 = all)")
    ap.add_argument("--batch-no-solo", action="store_true",
                    help="batch leg: skip the one-graph-at-a-time comparison (to profile the batched run alone)")
    ap.add_argument("--batch-max-vertices", type=int, default=0,
                    help="batch leg: vertices per launch group (0: planned from device memory by the library)")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-rate", type=float, default=200.0, help="serve leg: arrivals per second (Poisson)")
    ap.add_argument("--serve-min-n", type=int, default=1_000)
    ap.add_argument("--serve-max-n", type=int, default=2_000_000)
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--ooc-n", type=int, default=40_000_000)
    ap.add_argument("--ooc-partitions", type=int, default=64)
    ap.add_argument("--ooc-budget-gb", type=float, default=64.0)
    ap.add_argument("--ooc-check", action="store_true",
                    help="ooc mode: also close the graph in core (no budget) and compare checksums")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the closure engine runs on one GPU (one process, one device); --gpus must be 1"}))
        sys.exit(2)

    import graspan_g as gg
    from graspan_g import gen

    if args.mode == "batch":
        return print(json.dumps(line_of("graphs/sec (batched closure)", bench_batch(args), args)), flush=True)
    if args.mode == "serve":
        return print(json.dumps(line_of("graphs/sec (serving loop)", bench_serve(args), args)), flush=True)
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
    ctx.release()
    del ctx
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
        "roofline": roofline(stats),
        "breakdown_ms": {k: sum(s[k] for s in stats) / len(stats)
                         for k in ("fixpoint_ms", "expand_ms", "compact_ms", "merge_ms", "other_ms")},
        "join_candidates": s0["candidates"],
    }
    if not args.no_as_given:
        line["rule_table_as_given"] = as_given_leg(g, stats)
    if not args.no_batch:
        log("batched closure of the 4096-graph set")
        line["batch"] = bench_batch(args)
    if not args.no_serve:
        log(f"serving loop at {args.serve_rate} requests/s")
        line["serve"] = bench_serve(args)
    if not args.no_ooc:
        log(f"out-of-core closure of the {args.ooc_n}-vertex block dataflow graph")
        line["ooc"] = ooc_leg(args)
    if not args.no_cpu_baseline:
        log("cpu baseline: the oracle on the same n=%d graph" % g.n)
        line["cpu_baseline"] = cpu_baseline(g)
    print(json.dumps(line), flush=True)


def as_given_leg(g, planned_stats):
    """The same closure with the alias rule table evaluated as given
    (GG_STAR_REWRITE=0: V ::= L R joined directly), for comparison with the
    least-fixpoint rule plan the engine uses by default; one warm-up and two
    timed runs. The closure must be the same."""
    import graspan_g as gg

    os.environ["GG_STAR_REWRITE"] = "0"
    try:
        ctx = gg.Context()
        ctx.set_spec(g)
        ctx.run(0)
        t0 = time.perf_counter()
        sts = [ctx.run(0) for _ in range(2)]
        wall = time.perf_counter() - t0
        ctx.release()
    finally:
        del os.environ["GG_STAR_REWRITE"]
    st = sts[-1]
    if st.total_edges != planned_stats[-1]["total_edges"]:
        raise SystemExit("the rule table as given closes to a different edge count")
    log(f"rule table as given: {wall * 1e3 / 2:.1f} ms per closure, {st.rounds} rounds, "
        f"{st.candidates} join candidates")
    return {"ms_per_step": wall * 1e3 / 2, "value": 2 * (st.total_edges - st.input_unique) / wall,
            "rounds": st.rounds, "join_candidates": st.candidates,
            "note": "GG_STAR_REWRITE=0: the alias rule table evaluated as given (V ::= L R), same closure"}


def line_of(metric, leg, args):
    """A stand-alone JSON line for one leg (--mode batch / serve)."""
    return {"metric": metric, "value": leg["graphs_per_s"], "unit": "graphs/s", "n_gpus": 1, "steps": 1,
            "warmup": 0, "ms_per_step": leg["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u32", "data": "synthetic", "config": {"workload": leg["workload"]}, "detail": leg}


def bench_batch(args):
    """Batched closure of the 4096-graph set (vertex counts log-uniform in
    [500, 200000]; even graphs: dataflow grammar over the assign edges, odd
    graphs: alias grammar), in launch groups of at most
    --batch-max-vertices vertices, one launch sequence per fixpoint round and
    group; against the same kernels run one graph at a time. Every graph's
    batched result (per-label counts, checksum) must equal its solo result."""
    import numpy as np

    import graspan_g as gg
    from graspan_g import gen

    t = time.time()
    graphs = gen.batch_set(4096)[::args.batch_stride]
    nv = sum(g.n for g in graphs)
    log(f"generated {len(graphs)} graphs ({nv} vertices) in {time.time() - t:.1f}s")
    ctx = gg.Context()
    b = gg.Batch(ctx)
    for g in graphs:
        b.add_spec(g)
    t0 = time.perf_counter()
    st = b.run(max_vertices=args.batch_max_vertices)
    dt_b = time.perf_counter() - t0
    groups = b.info()[1]
    log(f"batched: {dt_b * 1e3:.0f} ms, {groups} launch groups, closure {st.total_edges}, "
        f"device peak {st.device_bytes_peak / 1e9:.1f} GB")
    batched = [b.result(i) for i in range(len(graphs))]
    b.close()
    ctx.release()
    if args.batch_no_solo:
        return {"workload": f"batched closure of {len(graphs)} graphs of the 4096-graph set "
                            f"(every {args.batch_stride}th), {nv} vertices, no one-at-a-time comparison",
                "graphs": len(graphs), "vertices": int(nv), "launch_groups": groups,
                "closure_edges": int(st.total_edges), "ms": dt_b * 1e3, "graphs_per_s": len(graphs) / dt_b,
                "closure_edges_per_s": (st.total_edges - st.input_unique) / dt_b,
                "device_bytes_peak": int(st.device_bytes_peak)}
    one = gg.Context()
    t0 = time.perf_counter()
    single = []
    solo_ms = []
    for g in graphs:
        t1 = time.perf_counter()
        one.set_spec(g)
        one.run(0)
        single.append((one.counts(), one.checksum()))
        solo_ms.append((time.perf_counter() - t1) * 1e3)
    dt_s = time.perf_counter() - t0
    one.release()
    log(f"one graph at a time: {dt_s * 1e3:.0f} ms")
    # where the one-at-a-time time goes, by graph size: small graphs are
    # bound by per-round costs that batching shares, large ones by the same
    # kernels either way
    by_size = []
    for lo, hi in ((0, 10_000), (10_000, 50_000), (50_000, 10**9)):
        idx = [i for i, g in enumerate(graphs) if lo <= g.n < hi]
        by_size.append({"n_from": lo, "n_to": hi if hi < 10**9 else None, "graphs": len(idx),
                        "one_at_a_time_ms": float(sum(solo_ms[i] for i in idx)),
                        "closure_edges": int(sum(int(single[i][0][1]) for i in idx))})
    for i, ((per, tot), cs) in enumerate(single):
        bp, bt, bc = batched[i]
        if (tot, cs) != (bt, bc) or not np.array_equal(per, bp):
            raise SystemExit(f"graph {i * args.batch_stride}: batched result differs from the single-graph result")
    return {"workload": f"batched closure of {len(graphs)} graphs of the 4096-graph set "
                        f"(every {args.batch_stride}th), {nv} vertices, launch groups "
                        + (f"of <= {args.batch_max_vertices} vertices" if args.batch_max_vertices
                           else "planned from device memory"),
            "graphs": len(graphs), "vertices": int(nv), "launch_groups": groups,
            "closure_edges": int(st.total_edges), "ms": dt_b * 1e3,
            "graphs_per_s": len(graphs) / dt_b, "one_at_a_time_ms": dt_s * 1e3,
            "one_at_a_time_graphs_per_s": len(graphs) / dt_s, "speedup": dt_s / dt_b,
            "one_at_a_time_by_size": by_size,
            "closure_edges_per_s": (st.total_edges - st.input_unique) / dt_b,
            "device_bytes_peak": int(st.device_bytes_peak),
            "checked": "every graph: batched per-label counts and checksum equal its one-at-a-time result"}


def bench_serve(args):
    """Serving loop under open-loop Poisson arrivals (--serve-rate requests/s,
    --serve-requests requests). Requests are alias-generator graphs with
    log-uniform vertex counts, grammar 50/50 dataflow / alias, admitted by
    estimated footprint under the budget. A seeded 1% sample (at least one
    request) is checked against the oracle."""
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
    res = [srv.wait(rid) for rid in ids]
    srv.close()
    lat = np.array([r.done_ms - r.submit_ms for r, _ in res])
    span = max(r.done_ms for r, _ in res) - min(r.submit_ms for r, _ in res)
    edges = sum(r.total_edges for r, _ in res)
    log(f"served {N} requests in {span:.0f} ms; p50 {np.percentile(lat, 50):.0f} ms, "
        f"p99 {np.percentile(lat, 99):.0f} ms")
    # seeded 1% sample against the oracle
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # checker only
    sample = sorted(np.random.default_rng(1).choice(N, size=max(1, N // 100), replace=False).tolist())
    for i in sample:
        ref = oracle.closure_of(graphs[i], want_triples=False, threads=host_threads())
        r, per = res[i]
        if (r.total_edges, r.checksum) != (ref.total, ref.checksum) or not np.array_equal(per, ref.counts):
            raise SystemExit(f"request {i}: served result differs from the oracle")
    log(f"oracle check of requests {sample}: equal")
    return {"workload": f"{N} requests, Poisson {args.serve_rate}/s, n log-uniform "
                        f"[{args.serve_min_n}, {args.serve_max_n}], grammar 50/50, {args.serve_slots} slots, "
                        f"{args.serve_budget_gb} GB budget",
            "requests": N, "offered_rate": args.serve_rate, "ms": span, "graphs_per_s": N / (span / 1e3),
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_batch_size": float(np.mean([r.batch_size for r, _ in res])),
            "closure_edges": int(edges), "closure_edges_per_s": edges / (span / 1e3),
            "oracle_checked_requests": sample,
            "oracle_checked_sizes": [int(sizes[i]) for i in sample]}


def ooc_leg(args):
    """Out-of-core closure: dataflow grammar over gen.dataflow_blocks (the
    40,000,000-vertex workload, ~1.2*10^8 e edges, ~7.8*10^9 closure edges),
    64 contiguous partitions, device budget 64 GB; partitions outside the
    current pair live in pinned host DRAM, and copies overlap the operators.
    The input starts on the host (the partitions are uploaded as the sweep
    needs them), so the time includes its first upload."""
    import graspan_g as gg
    from graspan_g import gen

    t = time.time()
    g = gen.dataflow_blocks(args.ooc_n, seed=0)
    log(f"generated blocks n={g.n} m={g.m} in {time.time() - t:.1f}s")
    ctx = gg.Context(device_budget=int(args.ooc_budget_gb * (1 << 30)), verbose=args.verbose or args.mode == "ooc")
    ctx.set_spec(g)
    t0 = time.perf_counter()
    st = ctx.run_out_of_core(args.ooc_partitions)
    wall = time.perf_counter() - t0
    per, total = ctx.counts()
    cs = ctx.checksum()
    log(f"out-of-core: {wall * 1e3:.0f} ms, closure {total}, pairs run {st.pairs_run} skipped {st.pairs_skipped}, "
        f"spilled {st.spill_bytes / 1e9:.1f} GB, device peak {st.device_bytes_peak / 1e9:.1f} GB, "
        f"{st.uploads_ahead} uploads and {st.spills_ahead} copy-outs ahead")
    ctx.release()
    del ctx
    check = None
    if args.ooc_check:
        inc = gg.Context()
        inc.set_spec(g)
        s2 = inc.run(0)
        check = {"in_core_total": int(s2.total_edges), "in_core_checksum_equal": inc.checksum() == cs,
                 "in_core_ms": s2.fixpoint_ms}
        if not check["in_core_checksum_equal"] or s2.total_edges != total:
            raise SystemExit("out-of-core closure differs from the in-core closure")
        log(f"in-core check: equal ({s2.fixpoint_ms:.0f} ms in core)")
    return {"metric": "closure edges/sec (out-of-core)", "value": (total - st.input_unique) / wall,
            "unit": "closure edges/s", "ms": wall * 1e3,
            "workload": f"dataflow blocks n={g.n}, {g.m} e edges, {args.ooc_partitions} partitions, "
                        f"{args.ooc_budget_gb} GB device budget",
            "closure_edges": int(total), "pairs_run": st.pairs_run, "pairs_skipped": st.pairs_skipped,
            "spill_bytes": int(st.spill_bytes), "device_bytes_peak": int(st.device_bytes_peak),
            "uploads_ahead": st.uploads_ahead, "spills_ahead": st.spills_ahead, "supersteps": st.supersteps,
            "check": check}


def bench_ooc(args):
    """--mode ooc: the out-of-core leg as its own line."""
    leg = ooc_leg(args)
    line = {
        "metric": leg["metric"],
        "value": leg["value"],
        "unit": leg["unit"],
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": leg["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": leg["workload"]},
        "detail": leg,
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
