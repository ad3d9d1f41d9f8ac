"""The allocation-free entry point (gg_closure_run_noalloc) against
gg_closure_run on the same graphs: per-label counts, checksum and every
closure edge, on two calls in a row (the second with a smaller input), with
the device's free memory unchanged by the calls; plus its refusals (a graph
too large for the fixed region, an edge out of range)."""
import ctypes
import sys

import numpy as np
import pytest

import graspan_g as gg
from graspan_g import gen

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, file=sys.stderr, flush=True)


def reference(g):
    c = gg.Context(device=0)
    c.set_spec(g)
    c.run(0)
    per, tot = c.counts()
    return per, tot, c.checksum(), c.export()


def csr_triples(row_ptr, col, n, L):
    lens = np.diff(row_ptr)
    rows = np.repeat(np.arange(L * n, dtype=np.int64), lens)
    src = (rows % n).astype(np.uint32)
    lab = (rows // n + 1).astype(np.uint32)
    t = np.stack([src, col.astype(np.uint32), lab], axis=1)
    order = np.lexsort((t[:, 2], t[:, 1], t[:, 0]))
    return t[order]


class Hip:
    """The few HIP runtime calls the test needs for its own device buffers,
    from the runtime the library itself is linked against (one runtime per
    process)."""

    def __init__(self):
        gg.lib()  # loads libamdhip64 as the library's dependency
        self.h = ctypes.CDLL("libamdhip64.so.7")
        P, S = ctypes.c_void_p, ctypes.c_size_t
        self.h.hipMalloc.argtypes = [ctypes.POINTER(P), S]
        self.h.hipFree.argtypes = [P]
        self.h.hipMemcpy.argtypes = [P, P, S, ctypes.c_int]
        self.h.hipMemset.argtypes = [P, ctypes.c_int, S]
        self.h.hipMemGetInfo.argtypes = [ctypes.POINTER(S), ctypes.POINTER(S)]

    def ok(self, rc, what):
        assert rc == 0, f"{what}: hip error {rc}"

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        self.ok(self.h.hipMalloc(ctypes.byref(p), max(nbytes, 4)), "hipMalloc")
        return p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.malloc(a.nbytes)
        if a.nbytes:
            self.ok(self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, p, count, dtype):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            self.ok(self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free_bytes(self):
        self.ok(self.h.hipDeviceSynchronize(), "hipDeviceSynchronize")
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        self.ok(self.h.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)), "hipMemGetInfo")
        return f.value


def noalloc(hip, g, col_cap):
    R = g.num_labels * g.n
    bufs = [hip.upload(g.src), hip.upload(g.dst), hip.upload(g.lab)]
    row_ptr = hip.malloc((R + 1) * 8)
    col = hip.malloc(max(col_cap, 1) * 4)
    bufs += [row_ptr, col]
    try:
        free0 = hip.free_bytes()
        total, cs, per = gg.closure_noalloc(g.rules, g.num_labels, g.n, g.m, bufs[0], bufs[1], bufs[2], row_ptr,
                                            col, col_cap, 0)
        free1 = hip.free_bytes()
        ptr = hip.download(row_ptr, R + 1, np.uint64).astype(np.int64)
        ids = hip.download(col, total, np.uint32)
        return total, cs, per, csr_triples(ptr, ids, g.n, g.num_labels), free0 - free1
    finally:
        for b in bufs:
            hip.h.hipFree(b)


def test_noalloc_matches_closure_run_twice(built):
    hip = Hip()
    warm, big, small = gen.alias(500, seed=300), gen.alias(3000, seed=301), gen.alias_as_dataflow(1500, seed=302)
    refs = [reference(g) for g in (warm, big, small)]  # also loads the library's kernels on the device
    # The first launch of a kernel that spills registers on a queue makes the
    # HIP runtime give that queue its private-segment memory; the first call
    # takes that, so the free-memory check covers the two calls after it.
    per, tot, cs, trip = refs[0]
    t2, cs2, _, trip2, _ = noalloc(hip, warm, tot)
    assert (t2, cs2) == (tot, cs)
    np.testing.assert_array_equal(trip2, trip)
    for g, (per, tot, cs, trip) in zip((big, small), refs[1:]):
        say("noalloc", g.name)
        t2, cs2, per2, trip2, used = noalloc(hip, g, tot + 10)
        assert (t2, cs2) == (tot, cs)
        np.testing.assert_array_equal(per2, per)
        np.testing.assert_array_equal(trip2, trip)
        assert used == 0, f"device free memory changed by {used} bytes"


def test_noalloc_refuses_what_does_not_fit(built):
    hip = Hip()
    g = gen.alias(40_000, seed=303)  # ~4e7 closure edges: more than the 256 MiB region holds
    say("noalloc too large", g.name)
    with pytest.raises(gg.GGError) as ei:
        noalloc(hip, g, 1 << 20)
    assert ei.value.status == gg.GG_E_NOMEM and "scratch" in str(ei.value)
    # an edge out of range
    bad = gen.GraphSpec(100, gen.DATAFLOW_RULES, np.array([1, 2], np.uint32), np.array([2, 100], np.uint32),
                        np.array([2, 2], np.uint8), "bad")
    with pytest.raises(gg.GGError) as ei:
        noalloc(hip, bad, 100)
    assert ei.value.status == gg.GG_E_ARG
    # the region is whole again for the next call
    ok = gen.alias(800, seed=304)
    per, tot, cs, trip = reference(ok)
    t2, cs2, _, trip2, _ = noalloc(hip, ok, tot)
    assert (t2, cs2) == (tot, cs)
    np.testing.assert_array_equal(trip2, trip)
