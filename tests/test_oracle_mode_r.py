"""Oracle mode R (the reference's partitioned schedule, needed for filter != 0)
against independent pins: with filter 0 it must reach the least fixpoint
(mode G and the golden vectors); with filter 1 it can only lose edges; the
dataflow grammar then derives nothing (its only route to N is the unary rule
that the filter skips); partitions follow initializePartitions."""
import numpy as np
import pytest

from conftest import golden_files, load_golden

import oracle
from graspan_g import gen


@pytest.mark.parametrize("path", golden_files(), ids=lambda p: p.rsplit("/", 1)[-1])
def test_mode_r_filter0_is_least_fixpoint(built, path):
    d = load_golden(path)
    r = oracle.closure_r(d["n"], d["rules_arr"], d["src"], d["dst"], d["lab"], d["num_labels"], 0)
    np.testing.assert_array_equal(r.triples, d["expected_arr"])


@pytest.mark.parametrize("maker", [lambda: gen.alias(700, seed=31), lambda: gen.dataflow(1500, seed=32),
                                   lambda: gen.alias_as_dataflow(4000, seed=33)])
def test_mode_r_filter0_equals_mode_g(built, maker):
    g = maker()
    a = oracle.closure_of(g)
    b = oracle.closure_r_of(g, filter_flag=0)
    assert b.checksum == a.checksum and b.total == a.total
    np.testing.assert_array_equal(b.counts, a.counts)
    assert len(b.parts) == 2 and b.parts[0][0] == 0 and b.parts[1][1] == g.n - 1


@pytest.mark.parametrize("maker", [lambda: gen.alias(1200, seed=34), lambda: gen.alias_as_dataflow(4000, seed=35)])
def test_mode_r_filter1_is_a_subset(built, maker):
    g = maker()
    full = oracle.closure_of(g)
    f1 = oracle.closure_r_of(g, filter_flag=1)
    t_full = set(map(tuple, full.triples.tolist()))
    t_f1 = set(map(tuple, f1.triples.tolist()))
    assert t_f1 <= t_full
    inp = set(zip(g.src.tolist(), g.dst.tolist(), g.lab.tolist()))
    assert inp <= t_f1


def test_mode_r_filter1_dataflow_derives_nothing(built):
    g = gen.dataflow(2000, seed=36)
    f1 = oracle.closure_r_of(g, filter_flag=1)
    assert f1.total == g.m  # only the input e edges: N ::= e is unary and skipped
    assert f1.counts[0] == 0


def test_mode_r_degenerate_split_rejected(built):
    """With no input edges the reference's small-graph split finds no second
    interval (its loop never passes the midpoint); the oracle refuses."""
    with pytest.raises(ValueError):
        oracle.closure_r(5, np.array([[1, 0, 0]], np.int32), np.zeros(0, np.uint32), np.zeros(0, np.uint32),
                         np.zeros(0, np.uint8), 1, 1)


def test_mode_r_partitions_split_at_element_midpoint(built):
    d = load_golden([p for p in golden_files() if p.endswith("sample.json")][0])
    r = oracle.closure_r(d["n"], d["rules_arr"], d["src"], d["dst"], d["lab"], d["num_labels"], 0)
    # elements of the sample: per source the distinct (dst/192, label) pairs
    elems = {}
    for s, t, l in zip(d["src"].tolist(), d["dst"].tolist(), d["lab"].tolist()):
        elems.setdefault(s, set()).add((t // 192, l))
    cnt = np.zeros(d["n"], np.int64)
    for s, e in elems.items():
        cnt[s] = len(e)
    total = cnt.sum()
    cut = int(np.nonzero(np.cumsum(cnt) > total // 2)[0][0])
    assert r.parts == [(0, cut - 1), (cut, d["n"] - 1)]
