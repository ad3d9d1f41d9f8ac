"""The least-fixpoint rule plan (gg_grammar_plan_fixpoint, host only).

A star label Y (rules exactly Y ::= eps and Y ::= Y P_k, or all Y ::= P_k Y)
is P*; a binary rule X ::= Y C whose lhs has no other rule is evaluated as
X ::= C plus X ::= P_k X (X ::= B Y as X ::= B plus X ::= X P_k). The plan
is checked here by what the mathematics fixes: the oracle, which runs the
rule table it is given, must reach the same closure (every triple) from the
planned table as from the original one, on the alias grammar and on random
grammars built around star labels; tables the rewrite must not touch come
back unchanged; a plan applied twice changes nothing more. The last test
plants plausible mistakes in a plan and checks that the closure comparison
catches them."""
import os

import numpy as np
import pytest

import graspan_g as gg
import oracle
from graspan_g import gen


def plan(rules, L):
    return gg.plan_fixpoint_rules(np.asarray(rules, np.int32).reshape(-1, 3), L)


def as_set(rules):
    return sorted(tuple(int(x) for x in r) for r in np.asarray(rules).reshape(-1, 3))


def closure_triples(n, rules, src, dst, lab, L):
    return oracle.closure(n, np.asarray(rules, np.int32).reshape(-1, 3), src, dst, lab, L).triples


def test_alias_grammar_plan():
    g = gen.alias(10, seed=0)
    ids = gen.label_ids(g.rule_lines)
    L, A, M, R, V = ids["L"], ids["A"], ids["M"], ids["R"], ids["V"]
    got = as_set(plan(g.rules, g.num_labels))
    want = [tuple(int(x) for x in r) for r in g.rules if tuple(r) != (V, L, R)]
    want += [(V, R, 0), (V, A, V), (V, M, V)]
    assert got == sorted(want)


def test_dataflow_grammar_unchanged():
    g = gen.dataflow(10, seed=0)
    assert as_set(plan(g.rules, g.num_labels)) == as_set(g.rules)


@pytest.mark.parametrize("rules,L", [
    # X (3) has a second rule
    ([[1, 0, 0], [1, 1, 2], [3, 1, 4], [3, 4, 0]], 4),
    # Y (1) has a unary rule, so it is not a star
    ([[1, 0, 0], [1, 1, 2], [1, 2, 0], [3, 1, 4]], 4),
    # Y has no eps rule
    ([[1, 1, 2], [3, 1, 4]], 4),
    # Y mixes left- and right-linear rules
    ([[1, 0, 0], [1, 1, 2], [1, 4, 1], [3, 1, 4]], 4),
    # a generator of Y is X itself
    ([[1, 0, 0], [1, 1, 3], [3, 1, 4]], 4),
    # X is one of its own operands
    ([[1, 0, 0], [1, 1, 2], [3, 1, 3]], 3),
])
def test_tables_left_alone(rules, L):
    assert as_set(plan(rules, L)) == as_set(rules)


def test_plan_is_idempotent_and_can_be_switched_off(monkeypatch):
    g = gen.alias(10, seed=0)
    p1 = plan(g.rules, g.num_labels)
    assert as_set(plan(p1, g.num_labels)) == as_set(p1)
    monkeypatch.setenv("GG_STAR_REWRITE", "0")
    assert as_set(plan(g.rules, g.num_labels)) == as_set(g.rules)


def random_star_grammar(rng, L=9):
    """Labels 1..3 terminals, 4..L nonterminals; 4 and 5 stars (left- and
    right-linear), the others built from random binary/unary rules, at least
    one of them a single binary rule over a star."""
    rules = []
    terms = [1, 2, 3]
    for y, left in ((4, True), (5, False)):
        rules.append([y, 0, 0])
        gens_ = rng.choice([1, 2, 3, 8, 9], size=rng.integers(1, 3), replace=False)
        for p in gens_:
            rules.append([y, y, int(p)] if left else [y, int(p), y])
    # X ::= Y C / X ::= B Y with single productions
    rules.append([6, 4, int(rng.choice([1, 2, 5, 7]))])
    rules.append([7, int(rng.choice([1, 3, 8])), 5])
    # nonterminals 8, 9 feed back into the stars' generators
    rules.append([8, int(rng.choice([6, 7])), int(rng.choice(terms))])
    rules.append([9, int(rng.choice(terms)), int(rng.choice([6, 7, 8]))])
    if rng.random() < 0.5:
        rules.append([9, int(rng.choice(terms)), 0])
    return np.asarray(rules, np.int32), L, terms


@pytest.mark.parametrize("seed", range(12))
def test_random_star_grammars_same_closure(built, seed):
    rng = np.random.default_rng(seed)
    rules, L, terms = random_star_grammar(rng)
    p = plan(rules, L)
    assert as_set(p) != as_set(rules)  # the rewrite applies
    n = int(rng.integers(20, 60))
    m = int(rng.integers(n, 3 * n))
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    lab = rng.choice(terms, m).astype(np.uint8)
    np.testing.assert_array_equal(closure_triples(n, p, src, dst, lab, L),
                                  closure_triples(n, rules, src, dst, lab, L))


def test_alias_graph_same_closure(built):
    g = gen.alias(3000, seed=5)
    p = plan(g.rules, g.num_labels)
    a = oracle.closure_of(g)
    b = oracle.closure(g.n, p, g.src, g.dst, g.lab, g.num_labels)
    np.testing.assert_array_equal(a.triples, b.triples)
    # and the planned table needs far fewer join candidates
    assert b.rounds[:, 1].sum() * 4 < a.rounds[:, 1].sum()


def test_planted_mistakes_are_caught(built):
    """Plausible mistakes in a plan change the closure, so the comparison
    above would fail on them."""
    g = gen.alias(2000, seed=6)
    ids = gen.label_ids(g.rule_lines)
    A, M, R, V, a = ids["A"], ids["M"], ids["R"], ids["V"], ids["a"]
    ref = oracle.closure_of(g).triples
    p = [tuple(int(x) for x in r) for r in plan(g.rules, g.num_labels)]
    mistakes = {
        "unary seed dropped": [r for r in p if r != (V, R, 0)],
        "one generator dropped": [r for r in p if r != (V, M, V)],
        "generators of the other star": [r for r in p if r not in ((V, A, V),)] + [(V, a, V)],
        "operands swapped": [r for r in p if r not in ((V, A, V), (V, M, V))] + [(V, V, A), (V, V, M)],
    }
    for what, rules in mistakes.items():
        got = closure_triples(g.n, rules, g.src, g.dst, g.lab, g.num_labels)
        assert got.shape != ref.shape or not np.array_equal(got, ref), what
