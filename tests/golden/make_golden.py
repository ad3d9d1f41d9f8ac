"""Writes the golden closure fixtures under tests/golden/.

Independent of the oracle: a naive (not semi-naive) fixpoint over Python sets,
written from the rule semantics the reference implements
(eps A: A(v,v) for every vertex, src/kernel.cu:445-464; unary A ::= B:
A(u,w) <= B(u,w), src/kernel.cu:422-442; binary A ::= B C:
A(u,w) <= B(u,v) & C(v,w), src/kernel.cu:297-362, 403-419), iterated until
nothing changes. Every rule is re-applied to the whole edge set each pass, so
it shares no code or structure with the oracle's delta bookkeeping.

Fixtures (JSON: n, rules, edges, expected sorted triples, per-label counts):
  tiny_*.json     seeded random graphs of <= 64 vertices under several grammars
  sample.json     the reference's own sample input (src/inputFiles/), closed
                  under its own grammar. The input edges are read from the
                  committed copy tests/golden/sample_graph.txt /
                  sample_grammar.txt, which hold the reference's data files.

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def parse_grammar(path):
    """Grammar file -> (rules as (lhs, rhs1, rhs2) with 0 = absent, names)."""
    lines = [l.rstrip("\n") for l in open(path)]
    data = [l for l in lines if not l.startswith("#")]
    nr = int(data[0])
    ids = {}
    rules = []
    for line in data[1:1 + nr]:
        items = line.split("\t")
        rule = [0, 0, 0]
        for j, it in enumerate(items):
            if it not in ids:
                ids[it] = len(ids) + 1
            rule[j] = ids[it]
        rules.append(tuple(rule))
    return rules, ids


def parse_graph(path, ids):
    lines = [l.rstrip("\n") for l in open(path)]
    data = [l for l in lines if not l.startswith("#")]
    m, n = (int(x) for x in data[0].split("\t"))
    edges = []
    for line in data[1:1 + m]:
        s, d, lab = line.split("\t")
        edges.append((int(s), int(d), ids[lab]))
    return n, edges


def naive_closure(n, rules, edges):
    """Least fixpoint by naive iteration over Python sets."""
    E = set(edges)
    for (a, b, c) in rules:
        if b == 0:
            E.update((v, v, a) for v in range(n))
    while True:
        by_label = {}
        for (s, d, l) in E:
            by_label.setdefault(l, {}).setdefault(s, set()).add(d)
        new = set()
        for (a, b, c) in rules:
            if b == 0:
                continue
            B = by_label.get(b, {})
            if c == 0:
                for s, ds in B.items():
                    new.update((s, d, a) for d in ds)
            else:
                C = by_label.get(c, {})
                for s, vs in B.items():
                    for v in vs:
                        for w in C.get(v, ()):
                            new.add((s, w, a))
        if new <= E:
            return sorted(E)
        E |= new


GRAMMARS = {
    "dataflow": [(1, 2, 0), (1, 1, 2)],                      # N ::= e | N e  (N=1, e=2)
    "alias": [(1, 0, 0), (1, 1, 2), (1, 1, 3), (4, 0, 0), (4, 4, 5), (4, 4, 3), (6, 1, 4), (7, 8, 6), (3, 7, 9)],
    # unary chains, a rule with identical rhs, a self-referential unary, and a
    # label that is never derived
    "mixed": [(1, 2, 0), (3, 1, 1), (4, 3, 0), (1, 4, 2), (5, 0, 0), (6, 5, 2), (6, 6, 6), (7, 7, 0)],
    # right-recursive dyck-like grammar with two bracket kinds
    "dyck": [(1, 0, 0), (2, 3, 1), (1, 2, 4), (5, 6, 1), (1, 5, 7), (1, 1, 1)],
}
NUM_LABELS = {"dataflow": 2, "alias": 9, "mixed": 7, "dyck": 7}
TERMINALS = {"dataflow": [2], "alias": [2, 5, 8, 9], "mixed": [2, 7], "dyck": [3, 4, 6, 7]}


def random_graph(rng, n, m, terminals, self_loops=True):
    edges = set()
    while len(edges) < m:
        s = rng.randrange(n)
        d = rng.randrange(n)
        if s == d and not self_loops:
            continue
        edges.add((s, d, rng.choice(terminals)))
    return sorted(edges)


def write(name, n, rules, L, edges, expected, extra=None):
    counts = [0] * L
    for (_, _, l) in expected:
        counts[l - 1] += 1
    doc = {"n": n, "num_labels": L, "rules": [list(r) for r in rules], "edges": [list(e) for e in edges],
           "expected": [list(t) for t in expected], "counts": counts}
    if extra:
        doc.update(extra)
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(doc, f, separators=(",", ":"))
    print(name, "n", n, "edges", len(edges), "closure", len(expected), file=sys.stderr)


def main():
    rng = random.Random(20241203)
    cases = [("dataflow", 12, 20), ("dataflow", 40, 60), ("alias", 16, 24), ("alias", 48, 90),
             ("mixed", 20, 30), ("mixed", 64, 80), ("dyck", 24, 40), ("dyck", 64, 100)]
    for i, (gname, n, m) in enumerate(cases):
        rules = GRAMMARS[gname]
        edges = random_graph(rng, n, m, TERMINALS[gname])
        write(f"tiny_{i:02d}_{gname}.json", n, rules, NUM_LABELS[gname], edges, naive_closure(n, rules, edges),
              {"grammar": gname})
    # duplicate edges and an isolated high vertex id: the input list repeats
    # edges (the closure keeps one copy)
    rules = GRAMMARS["dataflow"]
    edges = [(0, 1, 2), (0, 1, 2), (1, 2, 2), (2, 0, 2), (2, 0, 2), (5, 5, 2)]
    write("tiny_dups.json", 9, rules, 2, edges, naive_closure(9, rules, sorted(set(edges))), {"grammar": "dataflow"})
    # the reference's sample input under its own grammar
    rules, ids = parse_grammar(os.path.join(HERE, "sample_grammar.txt"))
    n, edges = parse_graph(os.path.join(HERE, "sample_graph.txt"), ids)
    names = [None] * len(ids)
    for k, v in ids.items():
        names[v - 1] = k
    write("sample.json", n, rules, len(ids), edges, naive_closure(n, rules, edges), {"label_names": names})


if __name__ == "__main__":
    main()
