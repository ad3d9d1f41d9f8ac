"""The CPU oracle against independent pins: golden brute-force fixtures, the
dataflow grammar's closed form (transitive closure), and invariants."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import breadth_first_order

from conftest import golden_files, load_golden

import oracle
from graspan_g import gen


@pytest.mark.parametrize("path", golden_files(), ids=lambda p: p.rsplit("/", 1)[-1])
def test_oracle_matches_golden(built, path):
    d = load_golden(path)
    r = oracle.closure(d["n"], d["rules_arr"], d["src"], d["dst"], d["lab"], d["num_labels"])
    assert r.total == len(d["expected"])
    np.testing.assert_array_equal(r.triples, d["expected_arr"])
    np.testing.assert_array_equal(r.counts, np.asarray(d["counts"], dtype=np.uint64))


def transitive_closure_pairs(n, src, dst):
    """Pairs (u, w) joined by a path of length >= 1 (scipy BFS per source)."""
    A = sp.csr_matrix((np.ones(len(src), np.int8), (src.astype(np.int64), dst.astype(np.int64))), shape=(n, n))
    out = []
    for u in range(n):
        # vertices reachable in >= 1 step = BFS from the successors of u
        succ = A.indices[A.indptr[u]:A.indptr[u + 1]]
        seen = set()
        for v in succ:
            if v not in seen:
                order = breadth_first_order(A, v, directed=True, return_predecessors=False)
                seen.update(order.tolist())
        out.extend((u, w) for w in sorted(seen))
    return np.asarray(out, dtype=np.uint32).reshape(-1, 2)


@pytest.mark.parametrize("n,window", [(200, 10), (600, 50)])
def test_oracle_dataflow_is_transitive_closure(built, n, window):
    g = gen.dataflow(n, seed=3, window=window)
    r = oracle.closure_of(g)
    N = r.triples[r.triples[:, 2] == 1][:, :2]
    e = r.triples[r.triples[:, 2] == 2][:, :2]
    np.testing.assert_array_equal(e, np.stack([g.src, g.dst], 1))
    np.testing.assert_array_equal(N, transitive_closure_pairs(n, g.src, g.dst))


def test_oracle_idempotent_and_monotone(built):
    g = gen.alias(3000, seed=5)
    r = oracle.closure_of(g)
    t = r.triples
    # input is contained in the closure
    inp = set(zip(g.src.tolist(), g.dst.tolist(), g.lab.tolist()))
    assert inp <= set(map(tuple, t.tolist()))
    # closing the closure adds nothing
    r2 = oracle.closure(g.n, g.rules, t[:, 0], t[:, 1], t[:, 2].astype(np.uint8), g.num_labels)
    np.testing.assert_array_equal(r2.triples, t)
    assert r2.rounds.shape[0] == 1 and r2.rounds[0, 3] == 0


def test_oracle_relabel_invariant(built):
    g = gen.alias(2000, seed=7)
    r = oracle.closure_of(g)
    perm = np.random.default_rng(1).permutation(g.n).astype(np.uint32)
    r2 = oracle.closure(g.n, g.rules, perm[g.src], perm[g.dst], g.lab, g.num_labels)
    t = r.triples.copy()
    t[:, 0] = perm[t[:, 0]]
    t[:, 1] = perm[t[:, 1]]
    t = t[np.lexsort((t[:, 2], t[:, 1], t[:, 0]))]
    np.testing.assert_array_equal(r2.triples, t)


def test_oracle_counters_consistent(built):
    g = gen.alias(3000, seed=9)
    r = oracle.closure_of(g, want_triples=False)
    P, K, O, N = r.rounds.sum(0)
    assert r.input_unique + N == r.total
    assert r.rounds[-1, 3] == 0            # last round finds nothing new
    assert P >= K >= N
    assert int(r.counts.sum()) == r.total


def test_golden_comparison_catches_a_dropped_rule(built):
    """A plausible mistake (one binary rule dropped) must change the closure, so
    the golden comparison is able to see it."""
    for path in golden_files():
        d = load_golden(path)
        rules = d["rules_arr"]
        binary = np.nonzero(rules[:, 2] != 0)[0]
        if len(binary) == 0:
            continue
        r = oracle.closure(d["n"], np.delete(rules, binary[-1], axis=0), d["src"], d["dst"], d["lab"],
                           d["num_labels"])
        if r.total != len(d["expected"]):
            return
    pytest.fail("no fixture is sensitive to dropping a binary rule")


def test_oracle_checksum_matches_triples(built):
    d = load_golden(golden_files()[0])
    r = oracle.closure(d["n"], d["rules_arr"], d["src"], d["dst"], d["lab"], d["num_labels"])
    from graspan_g import triple_hash  # product hash; both sides must agree on the definition
    s = sum(triple_hash(int(a), int(b), int(c)) for a, b, c in r.triples) % (1 << 64)
    assert s == r.checksum
