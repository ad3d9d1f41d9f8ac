"""Autotuner for the closure engine's launch and schedule parameters.

    python -m graspan_g.tune [--out PATH] [--passes 2]

Runs on the present GPU. Each trial sets every parameter through its
GG_TUNE_<NAME> environment variable (the engine reads them at the start of
every run) and times three workloads with the same kernels the library uses
(the fastest of a few runs after a warm-up):

  alias     the 1,000,000-vertex alias closure (3 runs);
  batch     every 32nd graph of the 4096-graph batched set, closed as a batch (2 runs);
  dataflow  the 10,000-vertex dataflow closure (5 runs).

The score of a setting is the sum over workloads of its time divided by the
time of the starting setting, so each workload weighs the same. Parameters
are tuned one at a time (coordinate descent over the candidate lists below),
`--passes` times. The result is written as "name value" lines, with the
measured times as comments, to tuned_gfx950.cfg beside libgraspan_g.so, where
the engine reads it when it starts.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import Batch, Context, PKG_DIR
from . import gen

# candidate values per parameter; the first entry of each list is not special
CANDIDATES = {
    "grab_light": [2, 4, 8],
    "grab_heavy": [1, 2, 4],
    "heavy_cand": [256, 1024, 4096, 16384],
    "guess_min": [2048, 4096, 8192],
    "defer_ratio": [32, 128, 512],
    "expand_bpc": [4, 5, 6, 7],
    "merge_rows_bpc": [4, 6, 8],
    "gather_bpc": [4, 6, 8],
    "merge_bpc": [4, 6, 8],
    "group_mem_share": [0.3, 0.4, 0.5],
    "light_cand": [0, 16, 32, 64],
    "light_cap": [16, 32, 64],
    "group_rows": [1, 2, 4, 8],
}
# the setting the previous tuning run chose (tuned_gfx950.cfg)
START = {"grab_light": 4, "grab_heavy": 1, "heavy_cand": 256, "guess_min": 4096, "defer_ratio": 512, "expand_bpc": 7,
         "merge_rows_bpc": 8, "gather_bpc": 6, "merge_bpc": 8, "group_mem_share": 0.4, "light_cand": 16,
         "light_cap": 16, "group_rows": 4}
# parameters that only the batched workload reads
BATCH_ONLY = {"group_mem_share"}


def log(*a):
    print("[tune]", *a, file=sys.stderr, flush=True)


class Workloads:
    def __init__(self):
        t = time.time()
        self.alias = gen.alias(1_000_000, seed=0)
        self.dataflow = gen.dataflow(10_000, seed=0)
        self.batch = gen.batch_set(4096)[::32]
        log(f"workloads generated in {time.time() - t:.1f}s")
        self.ctx = Context()

    def _best_of(self, reps: int, f) -> float:
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            f()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best

    def time_alias(self) -> float:
        self.ctx.set_spec(self.alias)
        self.ctx.run(0)
        return self._best_of(3, lambda: self.ctx.run(0))

    def time_dataflow(self) -> float:
        self.ctx.set_spec(self.dataflow)
        self.ctx.run(0)
        return self._best_of(5, lambda: self.ctx.run(0))

    def time_batch(self) -> float:
        b = Batch(self.ctx)
        for g in self.batch:
            b.add_spec(g)
        b.run()
        dt = self._best_of(2, lambda: b.run())
        b.close()
        return dt


def apply(setting: dict) -> None:
    for k, v in setting.items():
        os.environ["GG_TUNE_" + k.upper()] = str(v)


def measure(w: Workloads, setting: dict, only_batch: bool = False) -> dict:
    apply(setting)
    out = {"batch": w.time_batch()}
    if not only_batch:
        out["alias"] = w.time_alias()
        out["dataflow"] = w.time_dataflow()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=os.path.join(PKG_DIR, "tuned_gfx950.cfg"))
    ap.add_argument("--passes", type=int, default=1)
    args = ap.parse_args(argv)
    w = Workloads()
    best = dict(START)
    measure(w, best)  # first runs allocate and load: not measured
    base = measure(w, best)
    log("start", best, {k: f"{v * 1e3:.1f} ms" for k, v in base.items()})

    def score(m: dict) -> float:
        return sum(m[k] / base[k] for k in m)

    best_m = base
    best_s = score(base)
    trials = 1
    for p in range(args.passes):
        for name, values in CANDIDATES.items():
            for v in values:
                if v == best[name]:
                    continue
                trial = dict(best)
                trial[name] = v
                only_batch = name in BATCH_ONLY
                m = measure(w, trial, only_batch)
                trials += 1
                if only_batch:
                    m = dict(best_m, batch=m["batch"])
                s = score(m)
                log(f"pass {p} {name}={v}: " + ", ".join(f"{k} {x * 1e3:.1f} ms" for k, x in m.items()) +
                    f" score {s:.4f} (best {best_s:.4f})")
                # a change must win by more than the run-to-run noise
                if s < best_s - 0.003 * len(m):
                    best, best_m, best_s = trial, m, s
    with open(args.out, "w") as f:
        f.write("# Closure engine parameters picked by the autotuner on an MI355X:\n")
        f.write("#   python -m graspan_g.tune --out " + os.path.basename(args.out) + "\n")
        f.write(f"# {trials} settings measured; score = sum over workloads of time / starting time\n")
        for k in ("alias", "batch", "dataflow"):
            f.write(f"# {k}: start {base[k] * 1e3:.1f} ms, tuned {best_m[k] * 1e3:.1f} ms\n")
        for k in CANDIDATES:
            f.write(f"{k} {best[k]}\n")
    log("wrote", args.out, best)
    print(args.out)


if __name__ == "__main__":
    main()
