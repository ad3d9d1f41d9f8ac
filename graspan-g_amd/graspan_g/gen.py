"""Seeded synthetic program-graph generators for the closure workloads.

Every generator returns a :class:`GraphSpec` (vertex count, int32 rule
triples with 1-based label ids, and uint32 ``src``/``dst`` plus uint8 ``lab``
edge arrays with 1-based labels) and can be written out in the reference's
tab-separated file formats (src/main.cpp:69-97 for grammars,
src/main.cpp:106-139 and 177-221 for graphs) with :func:`write_files`.

Workloads
---------
dataflow(n, ...)  -- grammar ``N ::= e | N e`` over random ``e`` edges with a
                     locality window; the 10,000-vertex case is the small
                     plain-dataflow workload.
alias(n, ...)     -- pointer-alias program graph split into procedures with
                     assign ``a``/``A`` and deref ``d``/``D`` edges and call
                     sites to Zipf-distributed callees; the 1,000,000-vertex
                     case is the in-core alias workload.
batch_set(...)    -- many alias-generator graphs of log-uniform sizes, half
                     closed under the dataflow grammar and half under the
                     alias grammar.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Sequence, Tuple

import numpy as np

# Grammar rule lines exactly as they are written to a grammar file: one tuple of
# 1-3 label names per rule (eps, unary, binary).
DATAFLOW_RULES: Tuple[Tuple[str, ...], ...] = (("N", "e"), ("N", "N", "e"))
ALIAS_RULES: Tuple[Tuple[str, ...], ...] = (
    ("L",), ("L", "L", "A"), ("L", "L", "M"),
    ("R",), ("R", "R", "a"), ("R", "R", "M"),
    ("V", "L", "R"), ("T", "D", "V"), ("M", "T", "d"),
)


def label_ids(rule_lines: Sequence[Sequence[str]]) -> dict:
    """Label name -> id, assigned 1,2,... in order of first appearance, scanning
    each rule line left to right (src/main.cpp:77-88)."""
    ids: dict = {}
    for line in rule_lines:
        for name in line:
            if name not in ids:
                ids[name] = len(ids) + 1
    return ids


def rules_array(rule_lines: Sequence[Sequence[str]]) -> np.ndarray:
    """int32 [R,3] of (lhs, rhs1, rhs2) with 0 meaning absent (src/main.cpp:76-90)."""
    ids = label_ids(rule_lines)
    out = np.zeros((len(rule_lines), 3), dtype=np.int32)
    for i, line in enumerate(rule_lines):
        for j, name in enumerate(line):
            out[i, j] = ids[name]
    return out


@dataclasses.dataclass
class GraphSpec:
    n: int
    rule_lines: Tuple[Tuple[str, ...], ...]
    src: np.ndarray  # uint32
    dst: np.ndarray  # uint32
    lab: np.ndarray  # uint8, 1-based label id under rule_lines
    name: str = ""

    @property
    def rules(self) -> np.ndarray:
        return rules_array(self.rule_lines)

    @property
    def num_labels(self) -> int:
        return len(label_ids(self.rule_lines))

    @property
    def m(self) -> int:
        return int(self.src.shape[0])


def _dedup_sorted(src, dst, lab):
    """Drop duplicate (src,dst,lab) triples; order by (src, lab, dst) so the file
    is grouped by ascending src as the reference reader requires
    (src/main.cpp:125-129, 192-204)."""
    src = np.asarray(src, dtype=np.uint32)
    dst = np.asarray(dst, dtype=np.uint32)
    lab = np.asarray(lab, dtype=np.uint8)
    if src.size == 0 or int(src.max()) < (1 << 24):
        # (src, lab, dst) packed into 24 + 8 + 32 bits
        key = (src.astype(np.uint64) << np.uint64(40)) | (lab.astype(np.uint64) << np.uint64(32)) | dst.astype(np.uint64)
        key = np.unique(key)
        s = (key >> np.uint64(40)).astype(np.uint32)
        l = ((key >> np.uint64(32)) & np.uint64(0xFF)).astype(np.uint8)
        d = (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        return s, d, l
    if np.all(lab == lab[0]):
        # one label: (src, dst) in 32 + 32 bits
        key = np.unique((src.astype(np.uint64) << np.uint64(32)) | dst.astype(np.uint64))
        return ((key >> np.uint64(32)).astype(np.uint32), (key & np.uint64(0xFFFFFFFF)).astype(np.uint32),
                np.full(key.size, lab[0], dtype=np.uint8))
    # wider ids: (lab, dst) in 40 bits as the minor key, src the major key
    minor = (lab.astype(np.uint64) << np.uint64(32)) | dst.astype(np.uint64)
    order = np.lexsort((minor, src))
    s, minor = src[order], minor[order]
    keep = np.ones(s.size, dtype=bool)
    keep[1:] = (s[1:] != s[:-1]) | (minor[1:] != minor[:-1])
    s, minor = s[keep], minor[keep]
    return s, (minor & np.uint64(0xFFFFFFFF)).astype(np.uint32), (minor >> np.uint64(32)).astype(np.uint8)


def dataflow_edges(n: int, seed: int = 0, window: int = 50, p_local: float = 0.8,
                   p_geo: float = 1.0 / 3.0, max_deg: int = 16, rng=None):
    """Random ``e`` edges: each vertex i draws out-degree min(geometric(p), 16)
    (support >= 1); each edge goes to (i + U[1, window]) mod n with probability
    p_local, else to a uniform vertex. Self-loops and duplicates are dropped."""
    rng = np.random.default_rng(seed) if rng is None else rng
    deg = np.minimum(rng.geometric(p_geo, size=n), max_deg).astype(np.int64)
    src = np.repeat(np.arange(n, dtype=np.int64), deg)
    m = src.shape[0]
    local = rng.random(m) < p_local
    off = rng.integers(1, window + 1, size=m)
    uni = rng.integers(0, n, size=m)
    dst = np.where(local, (src + off) % n, uni)
    keep = dst != src
    return src[keep].astype(np.uint32), dst[keep].astype(np.uint32)


def dataflow(n: int = 10_000, seed: int = 0, window: int = 50) -> GraphSpec:
    """Plain dataflow workload: grammar N ::= e | N e over dataflow_edges."""
    src, dst = dataflow_edges(n, seed=seed, window=window)
    lab = np.full(src.shape[0], label_ids(DATAFLOW_RULES)["e"], dtype=np.uint8)
    s, d, l = _dedup_sorted(src, dst, lab)
    return GraphSpec(n, DATAFLOW_RULES, s, d, l, name=f"dataflow-n{n}-s{seed}")


def dataflow_blocks(n: int, seed: int = 0, window: int = 200, block_min: int = 150, block_max: int = 250,
                    p_geo: float = 1.0 / 3.0, max_deg: int = 16) -> GraphSpec:
    """Out-of-core dataflow workload (grammar N ::= e | N e).

    Interpretation of the workload text, fixed here: the small-workload
    generator with a locality window of `window` closes to ~n^2 edges (its
    wrap-around over all n vertices and its 20% uniform edges make one giant
    strongly connected component: 93.8% of the vertices at n = 2*10^4), not
    the "billions" of N edges the text expects at 4*10^7 vertices. Reach is
    therefore bounded: vertices form consecutive blocks of
    U[block_min, block_max] vertices (think procedures), each vertex draws
    min(geometric(p_geo), max_deg) out edges (support >= 1, ~3 per vertex as
    in the text) to (position + U[1, window]) mod block size inside its
    block, and there are no uniform long-range edges. Self-loops and
    duplicates are dropped. A block is then (nearly always) one strongly
    connected component, so the closure is about n * 200 N edges. Blocks
    ignore partition boundaries, so some straddle them and partition pairs
    have work."""
    rng = np.random.default_rng(seed)
    sizes = []
    tot = 0
    while tot < n:
        b = int(rng.integers(block_min, block_max + 1))
        b = min(b, n - tot)
        sizes.append(b)
        tot += b
    sizes = np.asarray(sizes, dtype=np.int64)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    bsize = np.repeat(sizes, sizes)
    bstart = np.repeat(starts, sizes)
    deg = np.minimum(rng.geometric(p_geo, size=n), max_deg).astype(np.int64)
    src = np.repeat(np.arange(n, dtype=np.int64), deg)
    pos = src - bstart[src]
    dst = bstart[src] + (pos + rng.integers(1, window + 1, size=src.shape[0])) % bsize[src]
    keep = dst != src
    lab = np.full(int(keep.sum()), label_ids(DATAFLOW_RULES)["e"], dtype=np.uint8)
    s, d, l = _dedup_sorted(src[keep].astype(np.uint32), dst[keep].astype(np.uint32), lab)
    return GraphSpec(n, DATAFLOW_RULES, s, d, l, name=f"dataflow-blocks-n{n}-s{seed}")


def _zipf_probs(k: int, a: float) -> np.ndarray:
    w = 1.0 / np.arange(1, k + 1, dtype=np.float64) ** a
    return w / w.sum()


def alias_edges(n: int, seed: int = 0, rng=None, proc_min: int = 20, proc_max: int = 200,
                lam: float = 1.5, p_deref: float = 0.1, zipf_a: float = 1.1, module_procs: int = 8):
    """Pointer-alias program graph (terminal edges only, as (src, dst, kind) with
    kind 0 = a, 1 = d).

    Interpretation of the workload text, fixed here:
    - procedures are contiguous vertex ranges of size U[proc_min, proc_max]
      (the last one is clipped to n);
    - vertex k of a procedure receives Poisson(lam) ``a`` edges, each from a
      uniform earlier vertex of the same procedure (vertex 0 receives none);
    - a p_deref fraction of vertices each get one ``d`` edge to a uniform other
      vertex of their procedure;
    - each procedure has U{1,3} call sites. Procedures are grouped into
      modules of `module_procs` consecutive procedures (compilation units);
      each procedure gets a random popularity rank within its module, and a
      call site picks its callee by Zipf(zipf_a) over those ranks among the
      procedures of the caller's module that come after the caller (an
      acyclic call graph in topological order, as in a program without
      recursion; the last procedure of a module makes no calls).
      module_procs=0 draws callees from all later procedures instead. With
      global callees the assign/deref flow percolates through the symmetric M
      relation and the closure grows as ~1.1 n^2 (measured at n = 3,000 and
      10,000), about 10^12 edges at 10^6 vertices; with modules of 8 it stays
      near 1,040 edges per vertex (measured at 10^4, 3*10^4, 10^5).
      The call site then adds U{2,4}
      ``a`` edges from uniform caller vertices (arguments) to the callee's first
      vertices (parameters 0..k-1, clipped to the callee size) and one ``a`` edge
      from the callee's last vertex (its return value) to a uniform caller
      vertex (the receiver).
    """
    rng = np.random.default_rng(seed) if rng is None else rng
    # procedure boundaries
    sizes = []
    tot = 0
    while tot < n:
        s = int(rng.integers(proc_min, proc_max + 1))
        s = min(s, n - tot)
        sizes.append(s)
        tot += s
    sizes = np.asarray(sizes, dtype=np.int64)
    P = sizes.shape[0]
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    proc_of = np.repeat(np.arange(P), sizes)
    pos = np.arange(n, dtype=np.int64) - starts[proc_of]

    # intra-procedure assign edges from earlier vertices
    cnt = rng.poisson(lam, size=n)
    cnt[pos == 0] = 0
    dst_a = np.repeat(np.arange(n, dtype=np.int64), cnt)
    pos_rep = np.repeat(pos, cnt)
    src_a = dst_a - pos_rep + (rng.random(dst_a.shape[0]) * pos_rep).astype(np.int64)

    # deref edges to another vertex of the same procedure
    has_d = (rng.random(n) < p_deref) & (sizes[proc_of] > 1)
    dv = np.nonzero(has_d)[0]
    psz = sizes[proc_of[dv]]
    r = (rng.random(dv.shape[0]) * (psz - 1)).astype(np.int64)
    tgt_pos = np.where(r >= pos[dv], r + 1, r)
    src_d = dv
    dst_d = starts[proc_of[dv]] + tgt_pos

    # call sites
    ncalls = rng.integers(1, 4, size=P)
    caller = np.repeat(np.arange(P), ncalls)
    if module_procs > 0:
        # popularity rank of each procedure within its module (1 = most popular)
        rank = np.empty(P, dtype=np.int64)
        for m0 in range(0, P, module_procs):
            m1 = min(m0 + module_procs, P)
            rank[m0:m1] = rng.permutation(m1 - m0) + 1
        weight = rank.astype(np.float64) ** (-zipf_a)
        callee = np.full(caller.shape[0], -1, dtype=np.int64)
        u = rng.random(caller.shape[0])
        for i, p in enumerate(caller.tolist()):
            hi = min((p // module_procs + 1) * module_procs, P)
            if p + 1 >= hi:
                continue
            cw = np.cumsum(weight[p + 1:hi])
            callee[i] = p + 1 + int(np.searchsorted(cw, u[i] * cw[-1], side="right"))
    else:
        rank_perm = rng.permutation(P)
        probs = _zipf_probs(P, zipf_a)
        callee = np.full(caller.shape[0], -1, dtype=np.int64)
        todo = np.nonzero(caller < P - 1)[0]
        for _ in range(64):  # rejection sampling restricted to later procedures
            if todo.shape[0] == 0:
                break
            pick = rank_perm[rng.choice(P, size=todo.shape[0], p=probs)]
            ok = pick > caller[todo]
            callee[todo[ok]] = pick[ok]
            todo = todo[~ok]
    keep_c = callee >= 0
    caller = caller[keep_c]
    callee = callee[keep_c]
    C = caller.shape[0]
    nparam = rng.integers(2, 5, size=C)
    nparam = np.minimum(nparam, sizes[callee])
    call_rep = np.repeat(np.arange(C), nparam)
    pidx = np.arange(call_rep.shape[0]) - np.repeat(np.cumsum(nparam) - nparam, nparam)
    cr = caller[call_rep]
    arg = starts[cr] + (rng.random(call_rep.shape[0]) * sizes[cr]).astype(np.int64)
    par = starts[callee[call_rep]] + pidx
    ret_src = starts[callee] + sizes[callee] - 1
    ret_dst = starts[caller] + (rng.random(C) * sizes[caller]).astype(np.int64)

    src = np.concatenate([src_a, arg, ret_src, src_d])
    dst = np.concatenate([dst_a, par, ret_dst, dst_d])
    kind = np.concatenate([np.zeros(src_a.shape[0] + arg.shape[0] + C, np.uint8),
                           np.ones(src_d.shape[0], np.uint8)])
    keep = src != dst
    return src[keep].astype(np.uint32), dst[keep].astype(np.uint32), kind[keep]


def alias(n: int = 1_000_000, seed: int = 0, **kw) -> GraphSpec:
    """In-core alias workload: alias_edges plus the reverse edges A (of a) and
    D (of d), under the alias grammar."""
    src, dst, kind = alias_edges(n, seed=seed, **kw)
    ids = label_ids(ALIAS_RULES)
    fwd = np.where(kind == 0, ids["a"], ids["d"]).astype(np.uint8)
    rev = np.where(kind == 0, ids["A"], ids["D"]).astype(np.uint8)
    s = np.concatenate([src, dst])
    d = np.concatenate([dst, src])
    l = np.concatenate([fwd, rev])
    s, d, l = _dedup_sorted(s, d, l)
    return GraphSpec(n, ALIAS_RULES, s, d, l, name=f"alias-n{n}-s{seed}")


def alias_as_dataflow(n: int, seed: int = 0, **kw) -> GraphSpec:
    """Dataflow grammar over the assign edges of an alias-generator graph
    (``e`` := ``a``)."""
    src, dst, kind = alias_edges(n, seed=seed, **kw)
    keep = kind == 0
    lab = np.full(int(keep.sum()), label_ids(DATAFLOW_RULES)["e"], dtype=np.uint8)
    s, d, l = _dedup_sorted(src[keep], dst[keep], lab)
    return GraphSpec(n, DATAFLOW_RULES, s, d, l, name=f"alias-df-n{n}-s{seed}")


def log_uniform_sizes(count: int, lo: int, hi: int, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size=count)).astype(np.int64)


def batch_set(count: int = 4096, lo: int = 500, hi: int = 200_000, seed: int = 0) -> List[GraphSpec]:
    """Batched workload: `count` graphs with log-uniform vertex counts; even
    indices use the dataflow grammar on the assign edges, odd ones the alias
    grammar."""
    sizes = log_uniform_sizes(count, lo, hi, seed)
    out = []
    for i, nv in enumerate(sizes):
        if i % 2 == 0:
            out.append(alias_as_dataflow(int(nv), seed=seed * 1_000_003 + i))
        else:
            out.append(alias(int(nv), seed=seed * 1_000_003 + i))
    return out


# ---------------------------------------------------------------------------
# reference-format writers

def write_grammar(path, rule_lines: Sequence[Sequence[str]]) -> None:
    with open(path, "w") as f:
        f.write("# number of rules\n%d\n# rules\n" % len(rule_lines))
        for line in rule_lines:
            f.write("\t".join(line) + "\n")


def write_graph(path, g: GraphSpec) -> None:
    names = {v: k for k, v in label_ids(g.rule_lines).items()}
    lab_names = np.array([names.get(i, "?") for i in range(256)], dtype=object)
    with open(path, "w") as f:
        f.write("# number of edges, number of variables\n%d\t%d\n# edges\n" % (g.m, g.n))
        chunk = 1 << 20
        for i in range(0, g.m, chunk):
            s = g.src[i:i + chunk]
            d = g.dst[i:i + chunk]
            l = lab_names[g.lab[i:i + chunk]]
            f.write("".join("%d\t%d\t%s\n" % t for t in zip(s.tolist(), d.tolist(), l.tolist())))


def write_files(g: GraphSpec, graph_path, grammar_path) -> None:
    write_graph(graph_path, g)
    write_grammar(grammar_path, g.rule_lines)
