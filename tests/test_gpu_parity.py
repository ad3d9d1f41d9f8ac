"""GPU closure (HIP engine through the C ABI) against the oracle and the
golden fixtures. Integer work: every comparison is bit-exact."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN, golden_files, load_golden

import graspan_g as gg
import oracle
from graspan_g import gen

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, file=sys.stderr, flush=True)


def gpu_closure(n, rules, src, dst, lab, L, **kw):
    c = gg.Context(**kw)
    c.set_grammar(rules, L)
    c.set_graph(n, src, dst, lab)
    st = c.run(0)
    return c, st


@pytest.mark.parametrize("path", golden_files(), ids=lambda p: p.rsplit("/", 1)[-1])
def test_gpu_matches_golden(built, path):
    d = load_golden(path)
    say("golden", d["name"])
    c, st = gpu_closure(d["n"], d["rules_arr"], d["src"], d["dst"], d["lab"], d["num_labels"])
    np.testing.assert_array_equal(c.export(), d["expected_arr"])
    per, tot = c.counts()
    assert tot == len(d["expected"]) == st.total_edges
    np.testing.assert_array_equal(per, np.asarray(d["counts"], np.uint64))


def check_against_oracle(g, full_triples=True, **ctx_kw):
    r = oracle.closure_of(g, want_triples=full_triples)
    c = gg.Context(**ctx_kw)
    c.set_spec(g)
    st = c.run(0)
    per, tot = c.counts()
    assert tot == r.total
    np.testing.assert_array_equal(per, r.counts)
    assert c.checksum() == r.checksum
    if full_triples:
        np.testing.assert_array_equal(c.export(), r.triples)
    # algorithmic counters and round count agree with the oracle's on the
    # rule table the engine evaluates (the least-fixpoint plan)
    planned = gg.plan_fixpoint_rules(g.rules, g.num_labels)
    rc = r
    if planned.shape != g.rules.shape or not np.array_equal(planned, g.rules):
        rc = oracle.closure(g.n, planned, g.src, g.dst, g.lab, g.num_labels, want_triples=False)
        assert rc.total == r.total and rc.checksum == r.checksum
    P, K, O, N = (int(x) for x in rc.rounds.sum(0))
    assert st.rounds == rc.rounds.shape[0]
    assert (st.probe_reads, st.candidates, st.old_reads, st.new_edges) == (P, K, O, N)
    assert st.input_unique == r.input_unique
    return c, st, r


def test_gpu_dataflow_small(built):
    say("dataflow n=2000")
    check_against_oracle(gen.dataflow(2000, seed=1))


def test_gpu_alias_small(built):
    say("alias n=5000")
    check_against_oracle(gen.alias(5000, seed=2))


def test_gpu_alias_rule_table_as_given(built, monkeypatch):
    """GG_STAR_REWRITE=0: the engine evaluates the alias rule table as given
    (V ::= L R joined directly); same closure, the oracle's own counters."""
    monkeypatch.setenv("GG_STAR_REWRITE", "0")
    say("alias n=5000, rule table as given")
    g = gen.alias(5000, seed=2)
    assert np.array_equal(gg.plan_fixpoint_rules(g.rules, g.num_labels), g.rules)
    check_against_oracle(g)


def test_gpu_dataflow_10k(built):
    """The small plain-dataflow workload (n = 10,000, filter 0)."""
    say("dataflow n=10000")
    check_against_oracle(gen.dataflow(10_000, seed=0), full_triples=False)


def test_gpu_alias_50k_wide_window(built):
    """n > 32768: rows get a computed candidate window."""
    say("alias n=50000")
    check_against_oracle(gen.alias(50_000, seed=3), full_triples=True)


def test_gpu_multichunk_rows(built):
    """Rows whose candidate window is wider than one LDS chunk (16384 ids)."""
    say("wide rows n=200000")
    rng = np.random.default_rng(4)
    n = 200_000
    # a few hubs with out-edges spread over the whole id range, chained
    src = np.concatenate([np.zeros(300, np.uint32), np.arange(1, 400, dtype=np.uint32)])
    dst = np.concatenate([rng.integers(0, n, 300).astype(np.uint32), rng.integers(0, n, 399).astype(np.uint32)])
    lab = np.full(src.shape[0], 2, np.uint8)
    g = gen.GraphSpec(n, gen.DATAFLOW_RULES, src, dst, lab, "wide")
    check_against_oracle(g)


def test_gpu_row_groups_and_their_deferrals(built, monkeypatch):
    """Rows expanded four per wave in one shared guessed window
    (k_expand_group): the closure and the P/K/O/N counters equal the
    oracle's, also when random far edges send rows to the wave kernel (a
    candidate outside the group's window), and GG_EXPAND_GROUP=0 (every row
    through the wave kernel) gives the same."""
    monkeypatch.setenv("GG_TUNE_LIGHT_CAND", "0")  # no lane-per-row rounds: every round groups its rows
    say("row groups: alias n=20000")
    g = gen.alias(20_000, seed=5)
    _, st, _ = check_against_oracle(g)
    assert st.group_rows > 0
    say("row groups: alias n=20000 plus far A edges")
    rng = np.random.default_rng(9)
    far = np.flatnonzero(rng.random(g.n) < 0.0005).astype(np.uint32)  # 7 edges: closure 2.5e7
    fdst = rng.integers(0, g.n, far.shape[0]).astype(np.uint32)
    a_lab = np.uint8(gen.label_ids(g.rule_lines)["A"])
    s, d, l = gen._dedup_sorted(np.concatenate([g.src, far]), np.concatenate([g.dst, fdst]),
                                np.concatenate([g.lab, np.full(far.shape[0], a_lab, np.uint8)]))
    gf = gen.GraphSpec(g.n, g.rule_lines, s, d, l, "alias-far-n20000")
    _, st2, r2 = check_against_oracle(gf)
    assert st2.group_rows > 0 and st2.group_deferred > 0
    monkeypatch.setenv("GG_EXPAND_GROUP", "0")
    say("row groups off: alias n=20000 plus far A edges")
    _, st3, r3 = check_against_oracle(gf, full_triples=False)
    assert st3.group_rows == 0 and r3.checksum == r2.checksum


def _paired_blocks(n, seed):
    """Block dataflow graph plus far edges u -> u + n/2 from 2% of the
    vertices of the lower half: a lower block reaches the blocks half the id
    range above it, so its rows hold two clusters of ids many LDS windows
    apart."""
    g = gen.dataflow_blocks(n, seed=seed)
    rng = np.random.default_rng(seed + 1000)
    far = np.flatnonzero(rng.random(n // 2) < 0.02).astype(np.uint32)
    src = np.concatenate([g.src, far])
    dst = np.concatenate([g.dst, far + np.uint32(n // 2)])
    s, d, l = gen._dedup_sorted(src, dst, np.full(src.shape[0], g.lab[0], np.uint8))
    return gen.GraphSpec(n, gen.DATAFLOW_RULES, s, d, l, f"paired-blocks-n{n}-s{seed}")


def test_gpu_sparse_wide_rows(built):
    """Rows whose ids sit in a few clusters far apart: the windows between
    the clusters are skipped, in core and in the partitioned schedule."""
    g = _paired_blocks(120_000, seed=61)
    say("sparse wide rows", g.name)
    c, st, ref = check_against_oracle(g, full_triples=False)
    o = gg.Context()
    o.set_spec(g)
    o.run_out_of_core(4)
    assert o.counts()[1] == ref.total and o.checksum() == ref.checksum


def test_gpu_staging_retry(built, monkeypatch):
    """A tiny first staging area forces the grow-and-retry path."""
    say("staging retry")
    monkeypatch.setenv("GG_STAGING_MIN_WORDS", "64")
    c, st, _ = check_against_oracle(gen.alias(3000, seed=6))
    assert st.retries > 0


def test_gpu_deterministic_and_rerunnable(built):
    say("rerun")
    g = gen.alias(8000, seed=8)
    c = gg.Context()
    c.set_spec(g)
    c.run(0)
    a = (c.counts()[1], c.checksum())
    c.run(0)
    assert (c.counts()[1], c.checksum()) == a


def test_gpu_empty_graph_eps_only(built):
    say("empty graph")
    c, st = gpu_closure(5, np.array([[1, 0, 0], [2, 1, 1]], np.int32), np.zeros(0, np.uint32),
                        np.zeros(0, np.uint32), np.zeros(0, np.uint8), 2)
    t = c.export()
    exp = [[v, v, l] for v in range(5) for l in (1, 2)]
    np.testing.assert_array_equal(t, np.asarray(exp, np.uint32))


def _cases_r():
    out = []
    for p in golden_files():
        d = load_golden(p)
        out.append((d["name"], d["n"], d["rules_arr"], d["src"], d["dst"], d["lab"], d["num_labels"]))
    for name, g in [("alias700", gen.alias(700, seed=41)), ("aliasdf2000", gen.alias_as_dataflow(2000, seed=42)),
                    ("dataflow800", gen.dataflow(800, seed=43))]:
        out.append((name, g.n, g.rules, g.src, g.dst, g.lab, g.num_labels))
    return out


@pytest.mark.parametrize("flt", [1, 0])
def test_gpu_partitioned_schedule_matches_oracle(built, flt):
    """The reference's partitioned schedule on the GPU against the oracle's
    transcription of it (mode R): same closure, same superstep count, same
    partitions; with filter 0 both also equal the least fixpoint."""
    for name, n, rules, src, dst, lab, L in _cases_r():
        say("partitioned", name, "filter", flt)
        ref = oracle.closure_r(n, rules, src, dst, lab, L, flt)
        c = gg.Context()
        c.set_grammar(rules, L)
        c.set_graph(n, src, dst, lab)
        assert c.reference_partitions() == ref.parts
        st = c.run(flt) if flt else c.run_partitioned(0)
        assert st.supersteps == ref.supersteps and st.partitions == len(ref.parts)
        np.testing.assert_array_equal(c.export(), ref.triples)
        if flt == 0:
            g_ref = oracle.closure(n, rules, src, dst, lab, L)
            np.testing.assert_array_equal(ref.triples, g_ref.triples)


def _fixture_arrays(f):
    e = np.asarray(f["edges"], np.int64).reshape(-1, 3)
    return (np.asarray(f["rules"], np.int32).reshape(-1, 3), e[:, 0].astype(np.uint32), e[:, 1].astype(np.uint32),
            e[:, 2].astype(np.uint8))


def test_gpu_filter_fixtures_match_hand_derivation(built, monkeypatch):
    """The hand-derived filter-1 closures (tests/golden/filter_fixtures.py):
    the GPU's partitioned schedule gives exactly the derived set, intervals
    and superstep count; with filter 0 every path gives the least fixpoint."""
    sys.path.insert(0, GOLDEN)
    from filter_fixtures import FIXTURES
    for f in FIXTURES:
        say("filter fixture", f["name"])
        if f["heap_words"]:
            monkeypatch.setenv("GG_REF_HEAP_WORDS", str(f["heap_words"]))
        else:
            monkeypatch.delenv("GG_REF_HEAP_WORDS", raising=False)
        rules, src, dst, lab = _fixture_arrays(f)
        c = gg.Context()
        c.set_grammar(rules, f["num_labels"])
        c.set_graph(f["n"], src, dst, lab)
        assert c.reference_partitions() == f["parts"]
        st = c.run(1)
        np.testing.assert_array_equal(c.export(), np.asarray(f["filter1"], np.uint32))
        assert st.supersteps == f["filter1_supersteps"] and st.partitions == len(f["parts"])
        want0 = np.asarray(f["filter0"], np.uint32)
        c.run_partitioned(0)
        np.testing.assert_array_equal(c.export(), want0)
        c.run(0)
        np.testing.assert_array_equal(c.export(), want0)


@pytest.mark.parametrize("flt", [1, 0])
def test_gpu_multi_partition_schedule_matches_oracle(built, monkeypatch, flt):
    """A small heap (GG_REF_HEAP_WORDS) makes initializePartitions' cut loop
    produce several intervals (src/main.cpp:145-155); the GPU's schedule over
    them equals mode R's: intervals, supersteps and the closure."""
    for name, g, parts_wanted in [("alias1500", gen.alias(1500, seed=81), 4),
                                  ("aliasdf4000", gen.alias_as_dataflow(4000, seed=82), 6),
                                  ("alias3000", gen.alias(3000, seed=83), 3)]:
        # elements per source as getGraphInfo counts them
        keys = (g.src.astype(np.uint64) << np.uint64(40)) | ((g.dst // 192).astype(np.uint64) << np.uint64(8)) | g.lab
        elems = np.unique(keys).shape[0]
        heap = 4 * (8 * elems // parts_wanted + ((g.n // parts_wanted + 31) // 32 * 32) * g.num_labels) + 4
        monkeypatch.setenv("GG_REF_HEAP_WORDS", str(heap))
        ref = oracle.closure_r_of(g, filter_flag=flt)
        say("multi-partition", name, "filter", flt, "intervals", len(ref.parts))
        assert len(ref.parts) >= 3
        c = gg.Context()
        c.set_spec(g)
        assert c.reference_partitions() == ref.parts
        st = c.run(flt) if flt else c.run_partitioned(0)
        assert st.supersteps == ref.supersteps and st.partitions == len(ref.parts)
        np.testing.assert_array_equal(c.export(), ref.triples)


def test_gpu_degrees_match_oracle(built):
    """gg_closure_degrees (the reference's per-vertex degree[old] column,
    getDegree / computeDegree, src/kernel.cu:782-839) equals the oracle's and
    the out-degree counted from the exported triples."""
    for g in (gen.alias(6000, seed=91), gen.alias_as_dataflow(9000, seed=92)):
        say("degrees", g.name)
        ref = oracle.closure_of(g, want_degrees=True)
        c = gg.Context()
        c.set_spec(g)
        c.run(0)
        deg = c.degrees()
        np.testing.assert_array_equal(deg, ref.degrees)
        np.testing.assert_array_equal(deg, np.bincount(c.export()[:, 0], minlength=g.n).astype(np.uint64))
        assert int(deg.sum()) == ref.total


def test_gpu_partitioned_degenerate_split_rejected(built):
    """No input edges: the reference's split has no second interval."""
    say("degenerate split")
    c = gg.Context()
    c.set_grammar(np.array([[1, 0, 0]], np.int32), 1)
    c.set_graph(5, np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint8))
    with pytest.raises(gg.GGError) as ei:
        c.run(1)
    assert ei.value.status == gg.GG_E_UNSUPPORTED


def test_comp_cli_report_and_dump(built, tmp_path):
    say("comp cli sample")
    dump = tmp_path / "out.bin"
    p = subprocess.run([gg.COMP_PATH, os.path.join(GOLDEN, "sample_graph.txt"),
                        os.path.join(GOLDEN, "sample_grammar.txt"), "0", "--dump", str(dump)],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    d = load_golden(os.path.join(GOLDEN, "sample.json"))
    lines = p.stdout.splitlines()
    for want in ["===== PREPROCESSING INFO =====", "NUM RULES: 12", "NUM LABELS: 13", "NUM VERTICES: 530",
                 "INITIAL NUM EDGES: 2000", "===== SPAGPU FINISHED =====",
                 f"FINAL NUM EDGES: {len(d['expected'])}"]:
        assert want in lines, want
    got = np.fromfile(dump, dtype=np.uint32).reshape(-1, 3)
    np.testing.assert_array_equal(got, d["expected_arr"])
    # filter 1: the reference's schedule, its partitions and superstep count
    ref = oracle.closure_r(d["n"], d["rules_arr"], d["src"], d["dst"], d["lab"], d["num_labels"], 1)
    p1 = subprocess.run([gg.COMP_PATH, os.path.join(GOLDEN, "sample_graph.txt"),
                         os.path.join(GOLDEN, "sample_grammar.txt"), "1"], capture_output=True, text=True, timeout=120)
    assert p1.returncode == 0, p1.stderr
    for want in [f"INITIAL NUM PARTITIONS: {len(ref.parts)}", f"NUM SUPERSTEPS: {ref.supersteps}",
                 f"FINAL NUM EDGES: {ref.total}"]:
        assert want in p1.stdout.splitlines(), want
    # code order (grammar first) is detected and swapped
    p2 = subprocess.run([gg.COMP_PATH, os.path.join(GOLDEN, "sample_grammar.txt"),
                         os.path.join(GOLDEN, "sample_graph.txt")], capture_output=True, text=True, timeout=120)
    assert p2.returncode == 0 and f"FINAL NUM EDGES: {len(d['expected'])}" in p2.stdout


def _partition_records(triples, parts):
    """Expected Partition<k> bytes: per interval, its closure edges by
    (src, label, dst), 5 bytes each (u32 dst little-endian, u8 label)."""
    out = []
    for a, b in parts:
        t = triples[(triples[:, 0] >= a) & (triples[:, 0] <= b)]
        t = t[np.lexsort((t[:, 1], t[:, 2], t[:, 0]))]
        rec = np.zeros((len(t), 5), np.uint8)
        rec[:, :4] = t[:, 1].astype("<u4").view(np.uint8).reshape(-1, 4)
        rec[:, 4] = t[:, 2]
        out.append(rec.tobytes())
    return out


def test_comp_partition_files(built, tmp_path):
    """--partition-files writes one Partition<k> file per interval of the
    reference's initializePartitions, holding the interval's closure edges
    in the record layout of its compressed partition store."""
    say("comp --partition-files sample")
    d = load_golden(os.path.join(GOLDEN, "sample.json"))
    out = tmp_path / "parts"
    out.mkdir()
    p = subprocess.run([gg.COMP_PATH, os.path.join(GOLDEN, "sample_graph.txt"),
                        os.path.join(GOLDEN, "sample_grammar.txt"), "0", "--partition-files", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    parts = oracle.partitions_r(d["n"], d["rules_arr"], d["src"], d["dst"], d["lab"], d["num_labels"])
    assert len(parts) == 2  # the sample takes the small-graph split
    want = _partition_records(np.asarray(d["expected_arr"]), parts)
    assert sorted(os.listdir(out)) == [f"Partition{k}" for k in range(len(parts))]
    for k, w in enumerate(want):
        assert (out / f"Partition{k}").read_bytes() == w, k
    assert sum(len(w) for w in want) == 5 * len(d["expected"])


def test_gpu_unpacked_path_matches(built, monkeypatch):
    """The same closure with packed old rows switched off (every candidate as a
    separate id)."""
    say("unpacked path alias n=20000")
    monkeypatch.setenv("GG_NO_PACK", "1")
    check_against_oracle(gen.alias(20_000, seed=12))


@pytest.mark.parametrize("ratio", ["1", "0"])
def test_gpu_deferred_merge_matches_oracle(built, monkeypatch, ratio):
    """Rounds' deltas kept in the pending CSR and read beside old (ratio 1:
    whenever pending + delta <= old, so most rounds defer) or merged into old
    every round (ratio 0). Both equal the oracle, P/K/O/N counters included."""
    monkeypatch.setenv("GG_DEFER_RATIO", ratio)
    for name, g in [("alias n=20000", gen.alias(20_000, seed=21)),
                    ("dataflow n=3000", gen.dataflow(3_000, seed=22)),
                    ("alias-as-dataflow n=30000", gen.alias_as_dataflow(30_000, seed=23))]:
        say("deferred merge ratio", ratio, name)
        check_against_oracle(g)
    say("deferred merge ratio", ratio, "unpacked alias n=20000")
    monkeypatch.setenv("GG_NO_PACK", "1")
    check_against_oracle(gen.alias(20_000, seed=24))


def test_gpu_packed_ids_kept_matches_oracle(built, monkeypatch):
    """The merges normally keep only the packed entries of packed labels'
    rows and write their ids after the last round; GG_KEEP_PACKED_IDS=1 keeps
    the ids in every merge. Both equal the oracle, wide rows included."""
    monkeypatch.setenv("GG_KEEP_PACKED_IDS", "1")
    say("packed ids kept: alias n=20000")
    check_against_oracle(gen.alias(20_000, seed=31))
    say("packed ids kept: wide rows n=120000")
    check_against_oracle(_paired_blocks(120_000, seed=32), full_triples=False)


def test_gpu_pair_arena_overflow_matches_oracle(built, monkeypatch):
    """GG_PAIR_ARENA=64 starts every round's pair-list pass with 64-word
    arenas: a wave reserves 1024 words at a time, so each pass with pairs
    overflows, skips its writes past the arenas and runs again with grown
    ones. The closure still equals the oracle, P/K/O/N counters included."""
    monkeypatch.setenv("GG_PAIR_ARENA", "64")
    # the dataflow grammar has no pair lists (its C label e is never derived)
    for name, g, has_pairs in [("alias n=20000", gen.alias(20_000, seed=41), True),
                               ("dataflow n=3000", gen.dataflow(3_000, seed=42), False)]:
        say("pair arenas forced small", name)
        _, st, _ = check_against_oracle(g)
        assert (st.pair_reruns > 0) == has_pairs  # the overflow-and-rerun path ran


def test_gpu_alias_1m_full_size(built):
    """The benchmark workload itself (alias generator, n = 1,000,000): per-label
    counts, checksum, round count and the P/K/O/N counters equal the oracle's."""
    say("alias n=1000000 (benchmark workload)")
    check_against_oracle(gen.alias(1_000_000, seed=0), full_triples=False)


def _mixed_set(count, lo, hi, seed):
    sizes = gen.log_uniform_sizes(count, lo, hi, seed)
    out = []
    for i, nv in enumerate(sizes):
        out.append(gen.alias_as_dataflow(int(nv), seed=seed * 7919 + i) if i % 2 == 0
                   else gen.alias(int(nv), seed=seed * 7919 + i))
    return out


def test_gpu_batch_matches_oracle(built):
    """Graphs of both grammars closed together equal their separate closures."""
    say("batch of 24 mixed graphs")
    graphs = _mixed_set(24, 50, 3000, seed=21)
    c = gg.Context()
    b = gg.Batch(c)
    ids = [b.add_spec(g) for g in graphs]
    b.run(keep_closures=True)
    assert b.info() == (24, 1)
    for gid, g in zip(ids, graphs):
        r = oracle.closure_of(g)
        per, tot, cs = b.result(gid)
        assert tot == r.total and cs == r.checksum
        np.testing.assert_array_equal(per, r.counts)
        np.testing.assert_array_equal(b.export(gid), r.triples)


def test_gpu_batch_groups_and_alone_equal(built):
    """Splitting the batch into several launch groups changes nothing, and a
    graph closed alone gives the same result as inside the batch."""
    say("batch split into groups")
    graphs = _mixed_set(16, 200, 5000, seed=22)
    c = gg.Context()
    b1 = gg.Batch(c)
    for g in graphs:
        b1.add_spec(g)
    b1.run()
    whole = [b1.result(i) for i in range(len(graphs))]
    b2 = gg.Batch(c)
    for g in graphs:
        b2.add_spec(g)
    b2.run(max_vertices=6000)
    assert b2.info()[1] > 1
    for i in range(len(graphs)):
        p1, t1, c1 = whole[i]
        p2, t2, c2 = b2.result(i)
        assert (t1, c1) == (t2, c2)
        np.testing.assert_array_equal(p1, p2)
    alone = gg.Context()
    alone.set_spec(graphs[3])
    alone.run(0)
    assert alone.counts()[1] == whole[3][1] and alone.checksum() == whole[3][2]


@pytest.mark.parametrize("lanes", ["1", "2"])
def test_gpu_batch_lanes_and_memory_splits(built, lanes, monkeypatch):
    """Launch groups run one at a time (one lane) or two at once (two
    engines on one device region); under a device budget far below the
    first planned group, groups run out of memory and are split in half
    until they fit. Every result equals the oracle's either way."""
    monkeypatch.setenv("GG_BATCH_LANES", lanes)
    graphs = _mixed_set(24, 5000, 20000, seed=24)
    # each engine may hold 1.25x the largest graph's own peak: every graph
    # fits alone, the whole set does not
    peak1 = 0
    for g in graphs:
        one = gg.Context()
        one.set_spec(g)
        peak1 = max(peak1, one.run(0).device_bytes_peak)
        one.release()
    budget = int(peak1 * 1.25) * int(lanes)
    say("batch lanes", lanes, "budget", budget)
    c = gg.Context(device_budget=budget)
    b = gg.Batch(c)
    ids = [b.add_spec(g) for g in graphs]
    b.run()
    assert b.info()[1] > 2
    for g, i in zip(graphs, ids):
        ref = oracle.closure_of(g, want_triples=False)
        per, tot, cs = b.result(i)
        assert (tot, cs) == (ref.total, ref.checksum), g.name
        np.testing.assert_array_equal(per, ref.counts)


def _relabelled(g, order, tag):
    """The same graph under its grammar's rule lines taken in another order:
    label ids follow first appearance, so the rule triples (and the batch's
    grammar key) differ while the closure is the same up to label names."""
    import dataclasses
    lines = tuple(g.rule_lines[i] for i in order)
    old_ids, new_ids = gen.label_ids(g.rule_lines), gen.label_ids(lines)
    lut = np.zeros(256, np.uint8)
    for name, i in old_ids.items():
        lut[i] = new_ids[name]
    return dataclasses.replace(g, rule_lines=lines, lab=lut[g.lab], name=f"{g.name}-{tag}")


def test_gpu_batch_many_distinct_grammars(built):
    """Twelve alias graphs under twelve distinct label numberings: 84 binary
    rules, more than one launch group may hold (64), so the batch must start
    new groups on the rule limit; every result equals the oracle's."""
    say("batch of 12 distinct alias grammars")
    rng = np.random.default_rng(23)
    graphs = []
    for k in range(12):
        base = gen.alias(int(rng.integers(200, 900)), seed=230 + k)
        graphs.append(_relabelled(base, rng.permutation(len(base.rule_lines)), f"p{k}"))
    assert len({tuple(map(tuple, g.rules.tolist())) for g in graphs}) == 12
    c = gg.Context()
    b = gg.Batch(c)
    ids = [b.add_spec(g) for g in graphs]
    b.run()
    assert b.info()[1] >= 2
    for g, i in zip(graphs, ids):
        ref = oracle.closure_of(g, want_triples=False)
        per, tot, cs = b.result(i)
        assert (tot, cs) == (ref.total, ref.checksum), g.name
        np.testing.assert_array_equal(per, ref.counts)


@pytest.mark.parametrize("keep_ids", ["0", "1"])
def test_gpu_server_results_equal_alone(built, keep_ids, monkeypatch):
    """Requests served in a mix (two slots, batches formed from whatever is
    queued) give the same closures as each request closed alone. Sessions
    keep packed labels' rows as packed entries only and read results from
    them (GG_KEEP_PACKED_IDS=0, the default) or keep their ids every round
    (1)."""
    monkeypatch.setenv("GG_KEEP_PACKED_IDS", keep_ids)
    say("server mix, keep ids", keep_ids)
    graphs = _mixed_set(40, 100, 20_000, seed=23)
    with gg.Server(slots=2, budget_bytes=2 << 30, bytes_per_vertex=16 << 10, max_batch_requests=8) as srv:
        ids = [srv.submit_spec(g) for g in graphs]
        results = [srv.wait(i) for i in ids]
    sizes = [r.batch_size for r, _ in results]
    assert max(sizes) > 1                       # requests were batched together
    assert len({r.slot for r, _ in results}) == 2  # both slots ran
    alone = gg.Context()
    for g, (r, per) in zip(graphs, results):
        assert r.status == gg.GG_OK
        alone.set_spec(g)
        alone.run(0)
        p2, t2 = alone.counts()
        assert (r.total_edges, r.checksum) == (t2, alone.checksum())
        np.testing.assert_array_equal(per, p2)
        assert r.submit_ms <= r.start_ms <= r.done_ms
        # and against the oracle, not only the engine's own solo run
        ref = oracle.closure_of(g, want_triples=False)
        assert (r.total_edges, r.checksum) == (ref.total, ref.checksum)
        np.testing.assert_array_equal(per, ref.counts)


def test_gpu_server_joins_running_fixpoint(built):
    """Continuous batching: with one slot, requests submitted while a large
    one is being closed join its running fixpoint at a round boundary and
    finish before it; every result equals the oracle's."""
    import time
    say("server: small requests join a running fixpoint")
    big = gen.alias(400_000, seed=24)
    small = [gen.alias(1500 + 100 * i, seed=25 + i) for i in range(6)]
    with gg.Server(slots=1, budget_bytes=64 << 30) as srv:
        rb = srv.submit_spec(big)
        time.sleep(0.02)
        rs = [srv.submit_spec(g) for g in small]
        res_small = [srv.wait(r) for r in rs]
        res_big = srv.wait(rb)
    b, _ = res_big
    joined = [r for r, _ in res_small if r.start_ms > b.start_ms and r.done_ms < b.done_ms]
    assert joined, "no small request joined the running fixpoint and finished first"
    for g, (r, per) in zip(small + [big], res_small + [res_big]):
        ref = oracle.closure_of(g, want_triples=False)
        assert (r.total_edges, r.checksum) == (ref.total, ref.checksum), g.name
        np.testing.assert_array_equal(per, ref.counts)


@pytest.mark.parametrize("parts", [1, 4, 16])
def test_gpu_out_of_core_matches_oracle(built, parts, monkeypatch):
    """Partitions spilled to host DRAM (at most two resident, through the
    GG_OOC_MAX_RESIDENT hook), pairs chosen by crossing edges: the least
    fixpoint, equal to the oracle's."""
    monkeypatch.setenv("GG_OOC_MAX_RESIDENT", "2")
    cases = [gen.dataflow_blocks(30_000, seed=51)]
    if parts <= 4:
        cases.append(gen.alias(1_000, seed=52))
    for g in cases:
        say("out-of-core", g.name, "partitions", parts)
        ref = oracle.closure_of(g, want_triples=False)
        c = gg.Context(device_budget=2 << 30)
        c.set_spec(g)
        st = c.run_out_of_core(parts)
        per, tot = c.counts()
        assert tot == ref.total and c.checksum() == ref.checksum
        np.testing.assert_array_equal(per, ref.counts)
        assert st.partitions == parts
        assert (st.spill_bytes > 0) == (parts > 2)
        if parts == 16 and g.name.startswith("dataflow"):
            # blocks only straddle neighbouring partitions
            assert st.pairs_skipped > st.pairs_run > 0


def test_gpu_out_of_core_budget_pressure(built):
    """With a device budget below the unconstrained peak, partitions leave
    HBM only when an allocation would not fit; the closure is unchanged."""
    g = gen.dataflow_blocks(400_000, seed=53)
    say("out-of-core budget", g.name)
    free = gg.Context()
    free.set_spec(g)
    s0 = free.run_out_of_core(16)
    ref = (free.counts()[1], free.checksum())
    assert s0.spill_bytes == 0
    assert s0.uploads_ahead > 0  # partitions come in during the step before their use
    free.release()
    budget = int(s0.device_bytes_peak * 0.75)
    c = gg.Context(device_budget=budget)
    c.set_spec(g)
    st = c.run_out_of_core(16)
    assert (c.counts()[1], c.checksum()) == ref
    assert st.spill_bytes > 0 and st.device_bytes_peak <= budget
    assert st.uploads_ahead + st.spills_ahead > 0
    inc = gg.Context()
    inc.set_spec(g)
    inc.run(0)
    assert (inc.counts()[1], inc.checksum()) == ref
