"""ctypes loader for the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's cpu_baseline leg. The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liborc.so")

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        lib.orc_run.restype = P
        lib.orc_run.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_int, P, ctypes.c_uint64, P, P, P,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.orc_error.restype = ctypes.c_char_p
        lib.orc_error.argtypes = [P]
        for name in ("orc_total", "orc_input_unique"):
            getattr(lib, name).restype = ctypes.c_uint64
            getattr(lib, name).argtypes = [P]
        lib.orc_counts.argtypes = [P, P]
        lib.orc_num_rounds.restype = ctypes.c_int
        lib.orc_num_rounds.argtypes = [P]
        lib.orc_round_stats.argtypes = [P, P]
        lib.orc_seconds.restype = ctypes.c_double
        lib.orc_seconds.argtypes = [P]
        lib.orc_degrees.argtypes = [P, P]
        lib.orc_export.restype = ctypes.c_uint64
        lib.orc_export.argtypes = [P, P, ctypes.c_uint64]
        lib.orc_free.argtypes = [P]
        lib.orc_checksum.restype = ctypes.c_uint64
        lib.orc_checksum.argtypes = [P]
        lib.orc_run_r.restype = P
        lib.orc_run_r.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_int, P, ctypes.c_uint64, P, P, P,
                                  ctypes.c_int]
        lib.orc_parts_r.restype = P
        lib.orc_parts_r.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_int, P, ctypes.c_uint64, P, P, P]
        lib.orc_supersteps.restype = ctypes.c_uint32
        lib.orc_supersteps.argtypes = [P]
        lib.orc_num_parts.restype = ctypes.c_uint32
        lib.orc_num_parts.argtypes = [P]
        lib.orc_parts.argtypes = [P, P, P]
        _lib = lib
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class OracleResult:
    n: int
    num_labels: int
    total: int
    input_unique: int
    counts: np.ndarray          # uint64 [L], label id l+1 at index l
    rounds: np.ndarray          # uint64 [R, 4] = (P, K, O, N)
    seconds: float
    triples: Optional[np.ndarray]  # uint32 [total, 3] sorted (src, dst, label)
    checksum: int = 0
    supersteps: int = 0          # mode R only
    parts: Optional[list] = None  # mode R only: [(first, last), ...]
    degrees: Optional[np.ndarray] = None  # uint64 [n]: closure out-degree per vertex over all labels

    @property
    def alg_bytes(self) -> int:
        """Sum over rounds of 4P + 8K + 4O + 8N."""
        r = self.rounds.astype(np.float64)
        return int((4 * r[:, 0] + 8 * r[:, 1] + 4 * r[:, 2] + 8 * r[:, 3]).sum())


def closure(n: int, rules: np.ndarray, src: np.ndarray, dst: np.ndarray, lab: np.ndarray,
            num_labels: int, filter_flag: int = 0, threads: int = 0, want_triples: bool = True,
            max_rounds: int = 0, want_degrees: bool = False) -> OracleResult:
    lib = _load()
    rules = np.ascontiguousarray(rules, dtype=np.int32).reshape(-1, 3)
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    lab = np.ascontiguousarray(lab, dtype=np.uint8)
    h = lib.orc_run(n, num_labels, rules.shape[0], _ptr(rules), src.shape[0], _ptr(src), _ptr(dst), _ptr(lab),
                    filter_flag, threads, max_rounds)
    try:
        err = lib.orc_error(h).decode()
        if err:
            raise ValueError(err)
        total = lib.orc_total(h)
        counts = np.zeros(num_labels, dtype=np.uint64)
        lib.orc_counts(h, _ptr(counts))
        nr = lib.orc_num_rounds(h)
        rounds = np.zeros((nr, 4), dtype=np.uint64)
        if nr:
            lib.orc_round_stats(h, _ptr(rounds))
        triples = None
        if want_triples:
            triples = np.empty((total, 3), dtype=np.uint32)
            lib.orc_export(h, _ptr(triples), total)
        degrees = None
        if want_degrees:
            # the reference's per-vertex edge count of old (getDegree,
            # src/kernel.cu:818-839) once the fixpoint is reached
            degrees = np.zeros(n, dtype=np.uint64)
            lib.orc_degrees(h, _ptr(degrees))
        return OracleResult(n, num_labels, int(total), int(lib.orc_input_unique(h)), counts, rounds,
                            float(lib.orc_seconds(h)), triples, int(lib.orc_checksum(h)), degrees=degrees)
    finally:
        lib.orc_free(h)


def closure_r(n: int, rules: np.ndarray, src: np.ndarray, dst: np.ndarray, lab: np.ndarray, num_labels: int,
              filter_flag: int = 0, want_triples: bool = True) -> OracleResult:
    """Mode R: the reference's partitioned schedule (any filter value)."""
    lib = _load()
    rules = np.ascontiguousarray(rules, dtype=np.int32).reshape(-1, 3)
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    lab = np.ascontiguousarray(lab, dtype=np.uint8)
    h = lib.orc_run_r(n, num_labels, rules.shape[0], _ptr(rules), src.shape[0], _ptr(src), _ptr(dst), _ptr(lab),
                      filter_flag)
    try:
        err = lib.orc_error(h).decode()
        if err:
            raise ValueError(err)
        total = lib.orc_total(h)
        counts = np.zeros(num_labels, dtype=np.uint64)
        lib.orc_counts(h, _ptr(counts))
        triples = None
        if want_triples:
            triples = np.empty((total, 3), dtype=np.uint32)
            lib.orc_export(h, _ptr(triples), total)
        npart = lib.orc_num_parts(h)
        first = np.zeros(npart, np.uint32)
        last = np.zeros(npart, np.uint32)
        lib.orc_parts(h, _ptr(first), _ptr(last))
        return OracleResult(n, num_labels, int(total), 0, counts, np.zeros((0, 4), np.uint64),
                            float(lib.orc_seconds(h)), triples, int(lib.orc_checksum(h)),
                            int(lib.orc_supersteps(h)), list(zip(first.tolist(), last.tolist())))
    finally:
        lib.orc_free(h)


def partitions_r(n: int, rules: np.ndarray, src: np.ndarray, dst: np.ndarray, lab: np.ndarray,
                 num_labels: int) -> list:
    """Mode R's vertex intervals [(first, last), ...] without running the
    closure; raises ValueError where the reference's cut is degenerate."""
    lib = _load()
    rules = np.ascontiguousarray(rules, dtype=np.int32).reshape(-1, 3)
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    lab = np.ascontiguousarray(lab, dtype=np.uint8)
    h = lib.orc_parts_r(n, num_labels, rules.shape[0], _ptr(rules), src.shape[0], _ptr(src), _ptr(dst), _ptr(lab))
    try:
        err = lib.orc_error(h).decode()
        if err:
            raise ValueError(err)
        npart = lib.orc_num_parts(h)
        first = np.zeros(npart, np.uint32)
        last = np.zeros(npart, np.uint32)
        lib.orc_parts(h, _ptr(first), _ptr(last))
        return list(zip(first.tolist(), last.tolist()))
    finally:
        lib.orc_free(h)


def closure_r_of(g, **kw) -> OracleResult:
    return closure_r(g.n, g.rules, g.src, g.dst, g.lab, g.num_labels, **kw)


def closure_of(g, **kw) -> OracleResult:
    """Closure of a generator GraphSpec."""
    return closure(g.n, g.rules, g.src, g.dst, g.lab, g.num_labels, **kw)
