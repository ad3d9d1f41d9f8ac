#!/usr/bin/env python3
"""Per-kernel hardware-counter summary of one rocprofv3 profile set.

    python3 summarize_counters.py DIR [--json OUT.json]

DIR holds the output of the four counter passes and the kernel trace listed
in profiles/expand_alias1m_counters.md (subfolders p_rd, p_wr, p_sq1, p_sq2
with <name>_counter_collection.csv, and p_trace with trace_kernel_stats.csv).
Prints a markdown table per kernel and, with --json, writes the figures that
bench.py reads for its roofline (the expand stage: every k_dedup_light,
k_expand_light, k_expand_group and k_expand launch together).

Conversions (MI355X, gfx950):
  HBM read bytes  = 32*TCC_EA0_RDREQ_32B + 64*TCC_EA0_RDREQ_64B + 128*TCC_EA0_RDREQ_128B
  HBM write bytes = 64*TCC_EA0_WRREQ_64B + 32*(TCC_EA0_WRREQ - TCC_EA0_WRREQ_64B)
  cycles          = GRBM_GUI_ACTIVE / 8 (reported once per XCC)
  VALU busy       = SQ_INSTS_VALU * 4 / (1024 SIMDs * cycles)
  waves per SIMD  = SQ_WAVE_CYCLES * 4 / (1024 * cycles)
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import os
import re
import sys

SIMDS = 1024
XCCS = 8
CUS = 256


def short(name: str) -> str:
    k = re.sub(r"^void ", "", name).replace("gg::(anonymous namespace)::", "")
    i = k.find("(")
    return k[:i] if i > 0 else k


def load(path):
    """counter name -> value summed over dispatches, per kernel; dispatch counts"""
    vals = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(set)
    with open(path) as f:
        for r in csv.DictReader(f):
            k = short(r["Kernel_Name"])
            vals[k][r["Counter_Name"]] += float(r["Counter_Value"])
            disp[k].add(r["Dispatch_Id"])
    return vals, {k: len(v) for k, v in disp.items()}


def find_csv(d, suffix):
    for root, _, files in os.walk(d):
        for f in files:
            if f.endswith(suffix):
                return os.path.join(root, f)
    raise SystemExit(f"no *{suffix} under {d}")


def summarize(d):
    passes = [load(find_csv(os.path.join(d, p), "_counter_collection.csv"))[0] for p in ("p_rd", "p_wr", "p_sq1")]
    sq2 = load(find_csv(os.path.join(d, "p_sq2"), "_counter_collection.csv"))[0]
    trace = {}
    with open(find_csv(os.path.join(d, "p_trace"), "_kernel_stats.csv")) as f:
        for r in csv.DictReader(f):
            trace[short(r["Name"])] = (int(r["Calls"]), float(r["TotalDurationNs"]) / 1e6)

    def group(keys):
        c = collections.defaultdict(float)
        for k in keys:
            for p in passes:
                for name, v in p.get(k, {}).items():
                    c[name] += v
            for name, v in sq2.get(k, {}).items():
                if name != "GRBM_GUI_ACTIVE":  # the sq1 pass carries the cycle count
                    c[name] += v
        calls = sum(trace.get(k, (0, 0.0))[0] for k in keys)
        ms = sum(trace.get(k, (0, 0.0))[1] for k in keys)
        rb = 32 * c["TCC_EA0_RDREQ_32B"] + 64 * c["TCC_EA0_RDREQ_64B"] + 128 * c["TCC_EA0_RDREQ_128B"]
        wb = 64 * c["TCC_EA0_WRREQ_64B"] + 32 * (c["TCC_EA0_WRREQ"] - c["TCC_EA0_WRREQ_64B"])
        cyc = c["GRBM_GUI_ACTIVE"] / XCCS
        simd_cycles = max(1.0, cyc * SIMDS)
        calls = max(1, calls)
        return {"launches": calls, "total_ms": ms, "avg_ms": ms / calls,
                "read_bytes_per_launch": rb / calls, "write_bytes_per_launch": wb / calls,
                "hbm_bytes_per_launch": (rb + wb) / calls, "hbm_gbs": (rb + wb) / max(1e-9, ms / 1e3) / 1e9,
                "valu_insts_per_launch": c["SQ_INSTS_VALU"] / calls,
                "salu_insts_per_launch": c["SQ_INSTS_SALU"] / calls,
                "lds_insts_per_launch": c["SQ_INSTS_LDS"] / calls,
                "clock_mhz": cyc / max(1e-9, ms / 1e3) / 1e6,
                "valu_busy_frac": c["SQ_INSTS_VALU"] * 4 / simd_cycles,
                "salu_per_valu": c["SQ_INSTS_SALU"] / max(1.0, c["SQ_INSTS_VALU"]),
                "lds_bank_conflict_cycles_per_cu_frac": c["SQ_LDS_BANK_CONFLICT"] / CUS / max(1.0, cyc),
                "mean_waves_per_simd": c["SQ_WAVE_CYCLES"] * 4 / simd_cycles}

    stage = sorted(k for k in trace if k.startswith("k_expand") or k.startswith("k_dedup_light"))
    res = {"expand stage (k_dedup_light + k_expand_light + k_expand_group + k_expand, all instances)": group(stage)}
    for k in stage + ["k_merge_rows", "k_pairs", "k_gather", "k_merge", "k_activate", "k_unpack"]:
        if k in trace:
            res[k] = group([k])
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("dir")
    ap.add_argument("--json", help="write the expand-stage figures bench.py reads")
    args = ap.parse_args()
    res = summarize(args.dir)
    print("| kernel | launches | avg ms | HBM GB/launch | HBM GB/s | % of 8 TB/s | VALU busy | SALU/VALU insts "
          "| waves/SIMD | LDS bank-conflict cycles |")
    print("|---|---|---|---|---|---|---|---|---|---|")
    for k, m in res.items():
        print(f"| {k} | {m['launches']} | {m['avg_ms']:.3f} | {m['hbm_bytes_per_launch'] / 1e9:.3f} | "
              f"{m['hbm_gbs']:.0f} | {m['hbm_gbs'] / 80:.1f} | {m['valu_busy_frac']:.2f} | {m['salu_per_valu']:.2f} | "
              f"{m['mean_waves_per_simd']:.2f} | {m['lds_bank_conflict_cycles_per_cu_frac']:.2f} |")
    if args.json:
        name, stage = next(iter(res.items()))
        out = {"kernel": name, "workload": "alias closure n=1,000,000 seed 0 (bench.py default, one step)",
               "cus": CUS}
        for key in ("hbm_bytes_per_launch", "read_bytes_per_launch", "write_bytes_per_launch",
                    "valu_insts_per_launch", "salu_insts_per_launch", "clock_mhz", "valu_busy_frac",
                    "mean_waves_per_simd", "lds_bank_conflict_cycles_per_cu_frac"):
            out[key] = stage[key]
        out["avg_ms_profiled"] = stage["avg_ms"]
        out["source"] = "profiles/expand_alias1m_counters.md"
        out["kernels"] = res
        with open(args.json, "w") as f:
            json.dump(out, f, indent=1)
        print("wrote", args.json, file=sys.stderr)


if __name__ == "__main__":
    main()
