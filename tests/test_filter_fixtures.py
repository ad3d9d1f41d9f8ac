"""Oracle mode R against the hand-derived filter_flag != 0 closures of
tests/golden/filter_fixtures.py: the exact filter-1 set, the superstep count
and the intervals, plus the least fixpoint with filter 0. Each fixture holds
an edge that a plausible transcription mistake would add or drop (an
old_B x delta_C join that must stay skipped, a tmp x tmp join inside spagpu
that must run)."""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

import graspan_g as gg
import oracle

sys.path.insert(0, GOLDEN)
from filter_fixtures import FIXTURES  # noqa: E402
from test_partitions import ref_partitions  # noqa: E402


def arrays(f):
    e = np.asarray(f["edges"], np.int64).reshape(-1, 3)
    return (np.asarray(f["rules"], np.int32).reshape(-1, 3), e[:, 0].astype(np.uint32), e[:, 1].astype(np.uint32),
            e[:, 2].astype(np.uint8))


def set_heap(f, monkeypatch):
    if f["heap_words"]:
        monkeypatch.setenv("GG_REF_HEAP_WORDS", str(f["heap_words"]))
    else:
        monkeypatch.delenv("GG_REF_HEAP_WORDS", raising=False)


@pytest.mark.parametrize("f", FIXTURES, ids=lambda f: f["name"])
def test_fixture_intervals(built, f, monkeypatch):
    set_heap(f, monkeypatch)
    rules, src, dst, lab = arrays(f)
    kw = {"heap_words": f["heap_words"]} if f["heap_words"] else {}
    assert ref_partitions(f["n"], f["num_labels"], src, dst, lab, **kw) == f["parts"]
    c = gg.Context()
    c.set_grammar(rules, f["num_labels"])
    c.set_graph(f["n"], src, dst, lab)
    assert c.reference_partitions() == f["parts"]


@pytest.mark.parametrize("f", FIXTURES, ids=lambda f: f["name"])
def test_mode_r_filter1_matches_hand_derivation(built, f, monkeypatch):
    set_heap(f, monkeypatch)
    rules, src, dst, lab = arrays(f)
    r = oracle.closure_r(f["n"], rules, src, dst, lab, f["num_labels"], 1)
    assert r.parts == f["parts"]
    np.testing.assert_array_equal(r.triples, np.asarray(f["filter1"], np.uint32))
    assert r.supersteps == f["filter1_supersteps"]


@pytest.mark.parametrize("f", FIXTURES, ids=lambda f: f["name"])
def test_filter0_least_fixpoint(built, f, monkeypatch):
    set_heap(f, monkeypatch)
    rules, src, dst, lab = arrays(f)
    want = np.asarray(f["filter0"], np.uint32)
    r = oracle.closure_r(f["n"], rules, src, dst, lab, f["num_labels"], 0)
    np.testing.assert_array_equal(r.triples, want)
    g = oracle.closure(f["n"], rules, src, dst, lab, f["num_labels"])
    np.testing.assert_array_equal(g.triples, want)


@pytest.mark.parametrize("mutation", [1, 2, 3])
def test_fixtures_catch_planted_mistakes(built, mutation, monkeypatch):
    """Each plausible transcription mistake, planted in the oracle's mode R,
    changes the filter-1 result of at least one hand-derived fixture:
    1 runs old_B x delta_C instead of delta_B x old_C, 2 drops spagpu's
    tmp x tmp launches, 3 runs spagpu's loop once."""
    monkeypatch.setenv("ORC_MODE_R_MUTATION", str(mutation))
    differs = []
    for f in FIXTURES:
        set_heap(f, monkeypatch)
        rules, src, dst, lab = arrays(f)
        r = oracle.closure_r(f["n"], rules, src, dst, lab, f["num_labels"], 1)
        if not np.array_equal(r.triples, np.asarray(f["filter1"], np.uint32)):
            differs.append(f["name"])
    assert differs, f"mutation {mutation} went unnoticed"
