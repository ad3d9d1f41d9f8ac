"""Python host mirror of the closure engine's C ABI (include/graspan_g.h).

A thin ctypes layer over the in-tree ``libgraspan_g.so``: it exposes the same
operations as the reference's driver (load grammar, load graph, run the
closure, report FINAL NUM EDGES; src/main.cpp:860-921) plus the sorted-triple
export, per-label counts and checksum used for parity. The compute path is
the HIP library; this module never computes a closure itself and raises if the
library is missing.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass
from typing import Optional

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libgraspan_g.so")
COMP_PATH = os.path.join(PKG_DIR, "comp")

GG_OK = 0
GG_E_ARG, GG_E_IO, GG_E_FORMAT, GG_E_STATE, GG_E_DEVICE, GG_E_NOMEM, GG_E_UNSUPPORTED = -1, -2, -3, -4, -5, -6, -7
_STATUS = {GG_E_ARG: "GG_E_ARG", GG_E_IO: "GG_E_IO", GG_E_FORMAT: "GG_E_FORMAT", GG_E_STATE: "GG_E_STATE",
           GG_E_DEVICE: "GG_E_DEVICE", GG_E_NOMEM: "GG_E_NOMEM", GG_E_UNSUPPORTED: "GG_E_UNSUPPORTED"}

# Every function declared in include/graspan_g.h.
EXPORTED = (
    "gg_ctx_create", "gg_ctx_destroy", "gg_last_error", "gg_version",
    "gg_grammar_load_file", "gg_grammar_set", "gg_grammar_info", "gg_grammar_rules", "gg_grammar_plan_fixpoint", "gg_label_name",
    "gg_label_id", "gg_graph_load_file", "gg_graph_set_coo", "gg_graph_info", "gg_graph_edges",
    "gg_closure_run", "gg_closure_run_partitioned", "gg_closure_run_out_of_core", "gg_reference_partitions", "gg_closure_counts", "gg_closure_export", "gg_closure_degrees", "gg_closure_checksum",
    "gg_triple_hash", "gg_closure_release", "gg_closure_run_noalloc", "gg_device_info",
    "gg_batch_create", "gg_batch_destroy", "gg_batch_add", "gg_batch_run", "gg_batch_info", "gg_batch_result",
    "gg_batch_export",
    "gg_server_create", "gg_server_submit", "gg_server_wait", "gg_server_now_ms", "gg_server_destroy",
)


class GGError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{_STATUS.get(status, status)}: {msg}")
        self.status = status


class Options(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("device_budget", ctypes.c_uint64), ("verbose", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [
        ("rounds", ctypes.c_uint32), ("retries", ctypes.c_uint32),
        ("input_unique", ctypes.c_uint64), ("total_edges", ctypes.c_uint64),
        ("fixpoint_ms", ctypes.c_double), ("expand_ms", ctypes.c_double), ("compact_ms", ctypes.c_double),
        ("merge_ms", ctypes.c_double), ("other_ms", ctypes.c_double),
        ("expand_launches", ctypes.c_uint32),
        ("probe_reads", ctypes.c_uint64), ("candidates", ctypes.c_uint64), ("old_reads", ctypes.c_uint64),
        ("new_edges", ctypes.c_uint64), ("device_bytes_peak", ctypes.c_uint64),
        ("supersteps", ctypes.c_uint32), ("partitions", ctypes.c_uint32),
        ("pairs_run", ctypes.c_uint32), ("pairs_skipped", ctypes.c_uint32), ("spill_bytes", ctypes.c_uint64),
        ("pair_reruns", ctypes.c_uint32), ("uploads_ahead", ctypes.c_uint32),
        ("spills_ahead", ctypes.c_uint32), ("pad0", ctypes.c_uint32), ("light_rows", ctypes.c_uint64),
        ("group_rows", ctypes.c_uint64), ("group_deferred", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def build() -> None:
    """Compiles libgraspan_g.so and comp in-tree (hipcc, gfx950)."""
    subprocess.run(["make", "-s", "-C", PKG_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run make -C {PKG_DIR}")
        L = ctypes.CDLL(LIB_PATH)
        P, I, U32, U64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64
        sig = {
            "gg_ctx_create": (I, [P, P]), "gg_ctx_destroy": (I, [P]), "gg_last_error": (ctypes.c_char_p, [P]),
            "gg_version": (ctypes.c_char_p, []),
            "gg_grammar_load_file": (I, [P, ctypes.c_char_p]), "gg_grammar_set": (I, [P, I, P, I]),
            "gg_grammar_info": (I, [P, P, P]), "gg_grammar_rules": (I, [P, P]),
            "gg_grammar_plan_fixpoint": (I, [I, P, I, P, I, P]),
            "gg_label_name": (ctypes.c_char_p, [P, I]), "gg_label_id": (I, [P, ctypes.c_char_p]),
            "gg_graph_load_file": (I, [P, ctypes.c_char_p]), "gg_graph_set_coo": (I, [P, U32, U64, P, P, P]),
            "gg_graph_info": (I, [P, P, P]), "gg_graph_edges": (I, [P, P, P, P]),
            "gg_closure_run": (I, [P, I, P]), "gg_closure_run_partitioned": (I, [P, I, P]),
            "gg_closure_run_out_of_core": (I, [P, I, P]),
            "gg_closure_counts": (I, [P, P, P]), "gg_reference_partitions": (I, [P, P, P, I, P]),
            "gg_closure_export": (I, [P, P, U64, P]), "gg_closure_degrees": (I, [P, P]),
            "gg_closure_checksum": (I, [P, P]), "gg_triple_hash": (U64, [U32, U32, U32]),
            "gg_closure_release": (I, [P]),
            "gg_closure_run_noalloc": (I, [I, P, I, U32, U64, P, P, P, P, P, P, U64, P, P, P, ctypes.c_char_p,
                                           ctypes.c_size_t]),
            "gg_device_info": (I, [P, P, P, ctypes.c_char_p, ctypes.c_size_t, P, P]),
            "gg_batch_create": (I, [P, P]), "gg_batch_destroy": (I, [P]),
            "gg_batch_add": (I, [P, I, P, I, U32, U64, P, P, P, P]),
            "gg_batch_run": (I, [P, U64, I, P]), "gg_batch_info": (I, [P, P, P]),
            "gg_batch_result": (I, [P, I, P, P, P]), "gg_batch_export": (I, [P, I, P, U64, P]),
            "gg_server_create": (I, [P, P]),
            "gg_server_submit": (I, [P, I, P, I, U32, U64, P, P, P, P]),
            "gg_server_wait": (I, [P, ctypes.c_int64, P, P]), "gg_server_now_ms": (ctypes.c_double, [P]),
            "gg_server_destroy": (I, [P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def triple_hash(s: int, d: int, l: int) -> int:
    return int(lib().gg_triple_hash(s, d, l))


def plan_fixpoint_rules(rules, num_labels: int) -> np.ndarray:
    """The rule triples the least-fixpoint engine evaluates for `rules`
    (gg_grammar_plan_fixpoint; host only): int32 [R', 3]."""
    rules = np.ascontiguousarray(rules, dtype=np.int32).reshape(-1, 3)
    cap = 4 * rules.shape[0] + 64
    out = np.zeros((cap, 3), dtype=np.int32)
    cnt = ctypes.c_int()
    st = lib().gg_grammar_plan_fixpoint(rules.shape[0], _ptr(rules), num_labels, _ptr(out), cap, ctypes.byref(cnt))
    if st != GG_OK:
        raise GGError(st, "gg_grammar_plan_fixpoint")
    return out[:cnt.value].copy()


@dataclass
class Closure:
    total: int
    counts: np.ndarray
    stats: dict


class Context:
    """One graph on one GPU (gg_ctx)."""

    def __init__(self, device: int = -1, device_budget: int = 0, verbose: bool = False):
        self._L = lib()
        self._h = ctypes.c_void_p()
        opts = Options(device, device_budget, 1 if verbose else 0)
        self._check(self._L.gg_ctx_create(ctypes.byref(opts), ctypes.byref(self._h)), "gg_ctx_create")

    def close(self):
        if self._h:
            self._L.gg_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, status: int, what: str):
        if status != GG_OK:
            msg = self._L.gg_last_error(self._h).decode() if self._h else ""
            raise GGError(status, f"{what}: {msg}")

    @property
    def last_error(self) -> str:
        return self._L.gg_last_error(self._h).decode()

    # grammar / graph
    def load_grammar(self, path: str):
        self._check(self._L.gg_grammar_load_file(self._h, os.fsencode(path)), "gg_grammar_load_file")

    def set_grammar(self, rules: np.ndarray, num_labels: int):
        rules = np.ascontiguousarray(rules, dtype=np.int32).reshape(-1, 3)
        self._check(self._L.gg_grammar_set(self._h, rules.shape[0], _ptr(rules), num_labels), "gg_grammar_set")

    def grammar_info(self):
        r, l = ctypes.c_int(), ctypes.c_int()
        self._check(self._L.gg_grammar_info(self._h, ctypes.byref(r), ctypes.byref(l)), "gg_grammar_info")
        return r.value, l.value

    def rules(self) -> np.ndarray:
        nr, _ = self.grammar_info()
        out = np.zeros((nr, 3), dtype=np.int32)
        self._check(self._L.gg_grammar_rules(self._h, _ptr(out)), "gg_grammar_rules")
        return out

    def label_name(self, label_id: int) -> Optional[str]:
        s = self._L.gg_label_name(self._h, label_id)
        return None if s is None else s.decode()

    def label_id(self, name: str) -> int:
        return int(self._L.gg_label_id(self._h, name.encode()))

    def load_graph(self, path: str):
        self._check(self._L.gg_graph_load_file(self._h, os.fsencode(path)), "gg_graph_load_file")

    def set_graph(self, n: int, src, dst, lab):
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        lab = np.ascontiguousarray(lab, dtype=np.uint8)
        self._check(self._L.gg_graph_set_coo(self._h, n, src.shape[0], _ptr(src), _ptr(dst), _ptr(lab)),
                    "gg_graph_set_coo")

    def set_spec(self, g):
        """Grammar + graph from a generator GraphSpec."""
        self.set_grammar(g.rules, g.num_labels)
        self.set_graph(g.n, g.src, g.dst, g.lab)

    def graph_info(self):
        n, m = ctypes.c_uint32(), ctypes.c_uint64()
        self._check(self._L.gg_graph_info(self._h, ctypes.byref(n), ctypes.byref(m)), "gg_graph_info")
        return n.value, m.value

    def graph_edges(self):
        _, m = self.graph_info()
        s = np.empty(m, np.uint32)
        d = np.empty(m, np.uint32)
        l = np.empty(m, np.uint8)
        self._check(self._L.gg_graph_edges(self._h, _ptr(s), _ptr(d), _ptr(l)), "gg_graph_edges")
        return s, d, l

    # closure
    def run(self, filter_flag: int = 0) -> Stats:
        st = Stats()
        self._check(self._L.gg_closure_run(self._h, filter_flag, ctypes.byref(st)), "gg_closure_run")
        return st

    def run_partitioned(self, filter_flag: int = 0) -> Stats:
        """The reference's partitioned schedule (any filter value)."""
        st = Stats()
        self._check(self._L.gg_closure_run_partitioned(self._h, filter_flag, ctypes.byref(st)),
                    "gg_closure_run_partitioned")
        return st

    def run_out_of_core(self, num_partitions: int) -> Stats:
        """Least fixpoint with partitions spilled to host DRAM (see the header)."""
        st = Stats()
        self._check(self._L.gg_closure_run_out_of_core(self._h, num_partitions, ctypes.byref(st)),
                    "gg_closure_run_out_of_core")
        return st

    def reference_partitions(self):
        cnt = ctypes.c_int()
        first = np.zeros(64, np.uint32)
        last = np.zeros(64, np.uint32)
        self._check(self._L.gg_reference_partitions(self._h, _ptr(first), _ptr(last), 64, ctypes.byref(cnt)),
                    "gg_reference_partitions")
        return list(zip(first[:cnt.value].tolist(), last[:cnt.value].tolist()))

    def counts(self):
        _, L = self.grammar_info()
        per = np.zeros(L, dtype=np.uint64)
        tot = ctypes.c_uint64()
        self._check(self._L.gg_closure_counts(self._h, _ptr(per), ctypes.byref(tot)), "gg_closure_counts")
        return per, tot.value

    def export(self) -> np.ndarray:
        cnt = ctypes.c_uint64()
        _, tot = self.counts()
        buf = np.empty((tot, 3), dtype=np.uint32)
        self._check(self._L.gg_closure_export(self._h, _ptr(buf), tot, ctypes.byref(cnt)), "gg_closure_export")
        return buf[: cnt.value]

    def degrees(self) -> np.ndarray:
        n, _ = self.graph_info()
        out = np.zeros(n, dtype=np.uint64)
        self._check(self._L.gg_closure_degrees(self._h, _ptr(out)), "gg_closure_degrees")
        return out

    def checksum(self) -> int:
        s = ctypes.c_uint64()
        self._check(self._L.gg_closure_checksum(self._h, ctypes.byref(s)), "gg_closure_checksum")
        return s.value

    def release(self):
        self._check(self._L.gg_closure_release(self._h), "gg_closure_release")

    def device_info(self) -> dict:
        cnt, sel = ctypes.c_int(), ctypes.c_int()
        name = ctypes.create_string_buffer(256)
        tot, fr = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self._L.gg_device_info(self._h, ctypes.byref(cnt), ctypes.byref(sel), name, 256,
                                           ctypes.byref(tot), ctypes.byref(fr)), "gg_device_info")
        return {"count": cnt.value, "selected": sel.value, "name": name.value.decode(), "total": tot.value,
                "free": fr.value}


class ServerOptions(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("slots", ctypes.c_int), ("budget_bytes", ctypes.c_uint64),
                ("bytes_per_vertex", ctypes.c_uint64), ("max_batch_requests", ctypes.c_uint32)]


class ServerResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("total_edges", ctypes.c_uint64), ("checksum", ctypes.c_uint64),
                ("submit_ms", ctypes.c_double), ("start_ms", ctypes.c_double), ("done_ms", ctypes.c_double),
                ("batch_size", ctypes.c_uint32), ("slot", ctypes.c_int)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Server:
    """Serving loop (gg_server_*): submit requests, wait for results."""

    def __init__(self, device: int = -1, slots: int = 2, budget_bytes: int = 200 << 30,
                 bytes_per_vertex: int = 32 << 10, max_batch_requests: int = 0):
        self._L = lib()
        self._h = ctypes.c_void_p()
        self._labels = {}
        o = ServerOptions(device, slots, budget_bytes, bytes_per_vertex, max_batch_requests)
        st = self._L.gg_server_create(ctypes.byref(o), ctypes.byref(self._h))
        if st != GG_OK:
            raise GGError(st, "gg_server_create")

    def submit(self, rules, num_labels, n, src, dst, lab) -> int:
        rules = np.ascontiguousarray(rules, dtype=np.int32).reshape(-1, 3)
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        lab = np.ascontiguousarray(lab, dtype=np.uint8)
        rid = ctypes.c_int64()
        st = self._L.gg_server_submit(self._h, rules.shape[0], _ptr(rules), num_labels, n, src.shape[0], _ptr(src),
                                      _ptr(dst), _ptr(lab), ctypes.byref(rid))
        if st != GG_OK:
            raise GGError(st, "gg_server_submit")
        self._labels[rid.value] = num_labels
        return rid.value

    def submit_spec(self, g) -> int:
        return self.submit(g.rules, g.num_labels, g.n, g.src, g.dst, g.lab)

    def wait(self, rid: int):
        res = ServerResult()
        per = np.zeros(self._labels[rid], dtype=np.uint64)
        st = self._L.gg_server_wait(self._h, rid, ctypes.byref(res), _ptr(per))
        if st != GG_OK:
            raise GGError(st, "gg_server_wait")
        if res.status != GG_OK:
            raise GGError(res.status, f"request {rid} failed")
        return res, per

    def now_ms(self) -> float:
        return float(self._L.gg_server_now_ms(self._h))

    def close(self):
        if self._h:
            self._L.gg_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many graphs closed together (gg_batch_*). Uses a Context's device."""

    def __init__(self, ctx: Context):
        self._ctx = ctx
        self._L = ctx._L
        self._h = ctypes.c_void_p()
        self._labels = []
        ctx._check(self._L.gg_batch_create(ctx._h, ctypes.byref(self._h)), "gg_batch_create")

    def close(self):
        if self._h:
            self._L.gg_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        if st != GG_OK:
            raise GGError(st, what)

    def add(self, rules, num_labels, n, src, dst, lab) -> int:
        rules = np.ascontiguousarray(rules, dtype=np.int32).reshape(-1, 3)
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        lab = np.ascontiguousarray(lab, dtype=np.uint8)
        gid = ctypes.c_int()
        self._check(self._L.gg_batch_add(self._h, rules.shape[0], _ptr(rules), num_labels, n, src.shape[0],
                                         _ptr(src), _ptr(dst), _ptr(lab), ctypes.byref(gid)), "gg_batch_add")
        self._labels.append(num_labels)
        return gid.value

    def add_spec(self, g) -> int:
        return self.add(g.rules, g.num_labels, g.n, g.src, g.dst, g.lab)

    def run(self, max_vertices: int = 0, keep_closures: bool = False) -> Stats:
        st = Stats()
        s = self._L.gg_batch_run(self._h, max_vertices, 1 if keep_closures else 0, ctypes.byref(st))
        if s != GG_OK:
            raise GGError(s, "gg_batch_run: " + self._ctx.last_error)
        return st

    def info(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        self._check(self._L.gg_batch_info(self._h, ctypes.byref(a), ctypes.byref(b)), "gg_batch_info")
        return a.value, b.value

    def result(self, gid: int):
        per = np.zeros(self._labels[gid], dtype=np.uint64)
        tot, cs = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self._L.gg_batch_result(self._h, gid, _ptr(per), ctypes.byref(tot), ctypes.byref(cs)),
                    "gg_batch_result")
        return per, tot.value, cs.value

    def export(self, gid: int) -> np.ndarray:
        _, tot, _ = self.result(gid)
        buf = np.empty((tot, 3), dtype=np.uint32)
        cnt = ctypes.c_uint64()
        self._check(self._L.gg_batch_export(self._h, gid, _ptr(buf), tot, ctypes.byref(cnt)), "gg_batch_export")
        return buf


def closure_noalloc(rules, num_labels: int, n: int, m: int, src_ptr: int, dst_ptr: int, lab_ptr: int,
                    row_ptr_ptr: int, col_ptr: int, col_cap: int, stream: int = 0):
    """gg_closure_run_noalloc on device buffers given by address (the input
    edges, the output CSR's row offsets and ids); runs on the current device
    and `stream`. Returns (total, checksum, per_label counts)."""
    L = lib()
    rules = np.ascontiguousarray(rules, dtype=np.int32).reshape(-1, 3)
    total, cs = ctypes.c_uint64(), ctypes.c_uint64()
    per = np.zeros(num_labels, dtype=np.uint64)
    err = ctypes.create_string_buffer(512)
    st = L.gg_closure_run_noalloc(rules.shape[0], _ptr(rules), num_labels, n, m, src_ptr, dst_ptr, lab_ptr,
                                  stream or None, row_ptr_ptr, col_ptr, col_cap, ctypes.byref(total),
                                  ctypes.byref(cs), _ptr(per), err, 512)
    if st != GG_OK:
        e = GGError(st, "gg_closure_run_noalloc: " + err.value.decode())
        e.total = total.value
        raise e
    return total.value, cs.value, per


def closure_of(g, filter_flag: int = 0, device: int = -1) -> tuple:
    """Closure of a generator GraphSpec on the GPU: (Context, Stats). The
    context keeps the closure for counts/export/checksum."""
    ctx = Context(device=device)
    ctx.set_spec(g)
    st = ctx.run(filter_flag)
    return ctx, st
