"""The reference's vertex intervals (initializePartitions, src/main.cpp:141-175,
over getGraphInfo's element pre-pass, src/main.cpp:106-139), restated here in
Python and compared with the library's gg_reference_partitions and with the
oracle's mode R. GG_REF_HEAP_WORDS shrinks the heap so that small graphs go
through the cut loop (src/main.cpp:145-155) and not only the small-graph split
(src/main.cpp:158-172)."""
import numpy as np
import pytest

from conftest import golden_files, load_golden

import graspan_g as gg
import oracle
from graspan_g import gen

REAL_HEAP = (10000 // 8 * 7) * 1024 * 256  # allocateElementPool, src/kernel.cu:35


def element_counts(n, src, dst, lab):
    """Per source, the number of distinct (dst / 192, label) pairs: the
    reference's 192-id bitmap elements (src/kernel.h:35-41), counted as
    getGraphInfo does on src-grouped input (src/main.cpp:120-135)."""
    order = np.argsort(src, kind="stable")
    cnt = [0] * n
    cur, seen = None, set()
    for k in order.tolist():
        s = int(src[k])
        if s != cur:
            if cur is not None:
                cnt[cur] = len(seen)
            cur, seen = s, set()
        seen.add((int(dst[k]) // 192, int(lab[k])))
    if cur is not None:
        cnt[cur] = len(seen)
    return cnt


def ref_partitions(n, L, src, dst, lab, heap_words=REAL_HEAP):
    """initializePartitions as written: returns the list of (first, last)
    intervals, or None where the reference's own arithmetic breaks (an empty
    interval from a cut at a vertex that alone overflows, or a small-graph
    split that finds no second interval)."""
    elems = element_counts(n, src, dst, lab)
    limit = heap_words // 4

    def rup32(x):
        return x if x % 32 == 0 else (x // 32 + 1) * 32

    parts, t1, t3 = [], 0, 0
    for i in range(n):
        t1 += elems[i]
        if t1 * 8 + rup32(i - t3 + 1) * L > limit:
            if i - 1 < t3:
                return None
            parts.append((t3, i - 1))
            t1, t3 = elems[i], i
    parts.append((t3, n - 1))
    if len(parts) == 1:
        t4 = 0
        for i in range(n):
            t4 += elems[i]
            if t4 > t1 // 2:
                if i == 0:
                    return None
                return [(0, i - 1), (i, n - 1)]
        return None
    return parts


def heap_for(n, L, elems_total, parts):
    """A heap size that gives roughly `parts` intervals."""
    per = 8 * elems_total // parts + ((n // parts + 31) // 32 * 32) * L
    return 4 * per + 4


def library_parts(n, rules, src, dst, lab, L):
    c = gg.Context()
    c.set_grammar(rules, L)
    c.set_graph(n, src, dst, lab)
    try:
        return c.reference_partitions()
    except gg.GGError as e:
        assert e.status == gg.GG_E_UNSUPPORTED
        return None


def oracle_parts(n, rules, src, dst, lab, L):
    try:
        return oracle.partitions_r(n, rules, src, dst, lab, L)
    except ValueError:
        return None


def _cases():
    out = []
    for p in golden_files():
        d = load_golden(p)
        out.append((d["name"], d["n"], d["rules_arr"], d["src"], d["dst"], d["lab"], d["num_labels"]))
    for g in (gen.alias(900, seed=71), gen.alias_as_dataflow(3000, seed=72), gen.dataflow(1200, seed=73)):
        out.append((g.name, g.n, g.rules, g.src, g.dst, g.lab, g.num_labels))
    # ids above 192: elements differ from out-label counts
    rng = np.random.default_rng(74)
    n = 2000
    src = rng.integers(0, n, 6000).astype(np.uint32)
    dst = rng.integers(0, n, 6000).astype(np.uint32)
    lab = rng.integers(1, 3, 6000).astype(np.uint8)
    out.append(("random2000", n, np.array([[1, 2, 0], [1, 1, 2]], np.int32), src, dst, lab, 2))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_partitions_match_restatement(built, case, monkeypatch):
    name, n, rules, src, dst, lab, L = case
    total = sum(element_counts(n, src, dst, lab))
    heaps = [None] + [heap_for(n, L, total, p) for p in (2, 3, 5, 9)]
    seen_multi = False
    for h in heaps:
        if h is None:
            monkeypatch.delenv("GG_REF_HEAP_WORDS", raising=False)
            want = ref_partitions(n, L, src, dst, lab)
        else:
            monkeypatch.setenv("GG_REF_HEAP_WORDS", str(h))
            want = ref_partitions(n, L, src, dst, lab, h)
        got_lib = library_parts(n, rules, src, dst, lab, L)
        got_orc = oracle_parts(n, rules, src, dst, lab, L)
        assert got_lib == want, (name, h)
        assert got_orc == want, (name, h)
        seen_multi |= want is not None and len(want) > 2
    assert seen_multi or n < 64  # the cut loop itself was reached


def test_partition_cut_degenerate_first_vertex_rejected(built, monkeypatch):
    """Vertex 0 alone passes the partition limit: the reference's cut would
    push the interval [0, 0 - 1] (an unsigned wrap, src/main.cpp:150). The
    restatement, the library and the oracle all refuse the graph."""
    n = 192 * 30
    src = np.zeros(20, np.uint32)
    dst = (np.arange(20, dtype=np.uint32) * 192 + 5).astype(np.uint32)  # 20 elements on vertex 0
    lab = np.ones(20, np.uint8)
    rules = np.array([[1, 1, 1]], np.int32)
    heap = 400  # limit 100 words < 8 * 20 + 32 * 1
    monkeypatch.setenv("GG_REF_HEAP_WORDS", str(heap))
    assert ref_partitions(n, 1, src, dst, lab, heap) is None
    assert library_parts(n, rules, src, dst, lab, 1) is None
    assert oracle_parts(n, rules, src, dst, lab, 1) is None
