#!/usr/bin/env python3
"""Known-size memory traffic for checking rocprofv3 HBM counters on MI355X.

Runs device kernels whose HBM bytes are known, so that the per-request-size
TCC counters (TCC_EA0_RDREQ_32B / _64B / _128B, TCC_EA0_WRREQ / _64B) can be
turned into bytes and checked before they are applied to the closure kernels
(the summaries under profiles/ give the commands):

  copy      a 1 GiB device-to-device copy: reads 1 GiB, writes 1 GiB;
  gather64  index_select of 2^22 random 64-byte rows out of a 1 GiB table;
  gather32  index_select of 2^22 random 32-byte rows out of a 1 GiB table.

Prints one JSON line with the bytes each phase must move. Measurement only:
the closure library does not use this file.
"""
import json
import sys


def main():
    import torch

    phases = set(sys.argv[1:]) or {"copy", "gather64", "gather32"}
    out = {}
    dev = "cuda"
    gib = 1 << 30
    if "copy" in phases:
        src = torch.ones(gib // 4, dtype=torch.int32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        out["copy"] = {"read_bytes": gib, "write_bytes": gib}
        del src, dst
    for name, width in (("gather64", 16), ("gather32", 8)):
        if name not in phases:
            continue
        rows = gib // (4 * width)
        table = torch.ones((rows, width), dtype=torch.int32, device=dev)
        g = torch.Generator(device=dev)
        g.manual_seed(0)
        k = 1 << 22
        idx = torch.randint(0, rows, (k,), device=dev, generator=g)
        torch.cuda.synchronize()
        res = torch.index_select(table, 0, idx)
        torch.cuda.synchronize()
        # distinct rows bound the table bytes a perfect cache would read
        distinct = int(torch.unique(idx).numel())
        out[name] = {"rows": k, "row_bytes": 4 * width, "distinct_rows": distinct,
                     "read_bytes_min": distinct * 4 * width + k * 8, "write_bytes": k * 4 * width}
        del table, idx, res
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
