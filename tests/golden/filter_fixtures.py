"""Hand-derived closures for filter_flag != 0 (fixture data, no code under test).

With a non-zero filter the reference skips every unary rule and every
old_B x delta_C join (src/kernel.cu:490,499,556,592,644,698,711,765), so its
result depends on the vertex intervals (initializePartitions,
src/main.cpp:141-175) and on the sweep (precomputation / run_computation,
src/main.cpp:764-858). Each fixture below was traced by hand through the
operator lists of spagpu_s (src/kernel.cu:466-521), spagpu_b (:523-628),
spagpu (:630-780) and mergeAndDiff (:943-985); the trace is written above the
fixture. Nothing here was produced by running the oracle or the engine.

Notation: X(u,v) is an edge u -> v with label X; "old", "D" (delta) and "tmp"
are the three regions of a partition; p0, p1, p2 are the intervals. A
compute11(src, dst1, dst2, tmp) launch joins the rows of `src` (keys v of
B(u, v)) with the C rows of v found in dst1 or dst2, each used only when v
lies in that region's interval; the result goes to the tmp of u's partition.

Elements (for the interval cut) are the distinct (dst / 192, label) pairs of
a source (getGraphInfo, src/main.cpp:106-139); every fixture keeps ids below
192, so a source's element count is its number of distinct out-labels.
"""

# ---------------------------------------------------------------------------
# F1: an edge reachable only through old_B x delta_C inside superstep 0.
#
# Grammar (labels by first appearance): A ::= B C (A=1, B=2, C=3);
# C ::= D E (D=4, E=5).
# Edges: B(0,1) D(1,2) E(2,3) and padding E(5,6) E(6,7) E(7,8) E(8,9)
# (E never starts a rule, so the padding only moves the cut).
# Elements: sources 0,1,2,5,6,7,8 hold one each, 7 in all. With the real heap
# one interval fits, so the small-graph split cuts at the first vertex whose
# running count passes 7/2 = 3: counts 1,2,3,3,3,4 -> vertex 5.
# Intervals: p0 = [0,4], p1 = [5,9].
#
# Superstep 0, p0 (old empty, D = B01 D12 E23):
#   iteration 1: compute11(D, D): B(0,1) has no C row at 1; D(1,2) x E(2,3)
#   -> tmp C(1,3). mergeAndDiff: old = B01 D12 E23, D = C13.
#   iteration 2: filter 1 runs only compute11(D, old, D): C(1,3) starts no
#   rule, nothing new. (With filter 0 compute11(old, D) would join
#   B(0,1) x C(1,3) -> A(0,3).) D empty: D <- old, old <- empty.
# Superstep 0, p1: E edges start no rule; D <- its input.
# Superstep 1, pair (0,1): spagpu_b runs compute11(D0, D1) and
#   compute11(D1, D0): every key of p0 (1, 2, 3) lies outside p1 and p1 holds
#   only E rows, so nothing; spagpu adds nothing (tmp stays empty).
#   mergeAndDiff moves D into old for both; every D is empty -> done.
# Result: the input plus C(1,3); A(0,3) is absent. NUM SUPERSTEPS: 1.
# The least fixpoint (filter 0) also holds A(0,3).
F1 = {
    "name": "old_x_delta_skipped_in_superstep0",
    "n": 10,
    "num_labels": 5,
    "rules": [[1, 2, 3], [3, 4, 5]],
    "edges": [[0, 1, 2], [1, 2, 4], [2, 3, 5], [5, 6, 5], [6, 7, 5], [7, 8, 5], [8, 9, 5]],
    "heap_words": None,
    "parts": [(0, 4), (5, 9)],
    "filter1": [[0, 1, 2], [1, 2, 4], [1, 3, 3], [2, 3, 5], [5, 6, 5], [6, 7, 5], [7, 8, 5], [8, 9, 5]],
    "filter1_supersteps": 1,
    "filter0": [[0, 1, 2], [0, 3, 1], [1, 2, 4], [1, 3, 3], [2, 3, 5], [5, 6, 5], [6, 7, 5], [7, 8, 5],
                [8, 9, 5]],
}

# ---------------------------------------------------------------------------
# F2: edges reachable only through tmp x tmp inside spagpu, and only in the
# second pass of spagpu's loop.
#
# Grammar: X ::= P Q (X=1, P=2, Q=3); Y ::= R S (Y=4, R=5, S=6);
# Z ::= X Y (Z=7); W ::= Z V (W=8, V=9).
# Edges: P(0,3) Q(3,1) R(1,4) S(4,2) V(2,4), and padding Q(4,0) (no P edge
# ends at 4, so it never joins; it moves the cut).
# Elements: sources 0,1,2,3 one each, source 4 two (S, Q): 6 in all; the
# split cuts at the first vertex whose running count passes 3:
# counts 1,2,3,4 -> vertex 3. Intervals: p0 = [0,2], p1 = [3,4].
#
# Superstep 0: p0's keys 3 and 4 lie outside p0, V starts no rule, p1's rows
# (Q, S) start no rule; nothing is derived. Both: D <- input rows, old empty.
# Superstep 1, pair (0,1):
#   spagpu_b: compute11(D0, D1): P(0,3) x Q(3,1) -> tmp0 X(0,1);
#   R(1,4) x S(4,2) -> tmp0 Y(1,2). compute11(D1, D0): nothing.
#   spagpu, iteration 1, p0 side:
#     compute11(D0, tmp0, tmp1): the C rows it needs (Q at 3, S at 4) would be
#     in tmp1, which is empty;
#     compute11(tmp0, D0): X(0,1) needs Y at 1 in D0 -- Y(1,2) is in tmp0;
#     compute11(tmp0, D1): key 1 is outside p1;
#     compute11(tmp0, tmp0, tmp1): X(0,1) x Y(1,2) -> tmp0 Z(0,2): changed.
#   iteration 2, p0 side: compute11(tmp0, D0): Z(0,2) x V(2,4) -> tmp0 W(0,4):
#   changed. Iteration 3 adds nothing.
#   mergeAndDiff(p0): old0 = P03 R14 V24, D0 = X01 Y12 Z02 W04.
#   mergeAndDiff(p1): old1 = Q31 S42 Q40, D1 empty.
# Superstep 2, pair (0,1): compute11(D0, old1): keys 1 and 2 are outside p1,
#   W starts no rule; spagpu's tmp stays empty. Intra-partition D0 x D0 and
#   D0 x old0 are never run after superstep 0, so without the tmp x tmp launch
#   Z(0,2) and W(0,4) would be missing, and with one pass of spagpu's loop
#   W(0,4) would be. mergeAndDiff(p0): D0 into old0, every D empty -> done.
# Result: the input plus X(0,1), Y(1,2), Z(0,2), W(0,4). NUM SUPERSTEPS: 2.
# The least fixpoint is the same set.
F2 = {
    "name": "tmp_x_tmp_in_spagpu",
    "n": 5,
    "num_labels": 9,
    "rules": [[1, 2, 3], [4, 5, 6], [7, 1, 4], [8, 7, 9]],
    "edges": [[0, 3, 2], [3, 1, 3], [1, 4, 5], [4, 2, 6], [2, 4, 9], [4, 0, 3]],
    "heap_words": None,
    "parts": [(0, 2), (3, 4)],
    "filter1": [[0, 1, 1], [0, 2, 7], [0, 3, 2], [0, 4, 8], [1, 2, 4], [1, 4, 5], [2, 4, 9], [3, 1, 3],
                [4, 0, 3], [4, 2, 6]],
    "filter1_supersteps": 2,
    "filter0": [[0, 1, 1], [0, 2, 7], [0, 3, 2], [0, 4, 8], [1, 2, 4], [1, 4, 5], [2, 4, 9], [3, 1, 3],
                [4, 0, 3], [4, 2, 6]],
}

# ---------------------------------------------------------------------------
# F3: old_B x delta_C across two partitions in spagpu_b, skipped; needs three
# intervals, so the heap is set small enough for the cut loop of
# src/main.cpp:145-155.
#
# Grammar: A ::= B C (A=1, B=2, C=3); C ::= D E (D=4, E=5). L = 5.
# Edges: B(0,2) D(2,4) E(4,5); n = 6.
# Heap 680 words -> a partition may hold 680/4 = 170 words of
# 8 * elements + roundup32(vertices) * L. Running the loop:
#   i=0: 8*1 + 32*5 = 168; i=1: 168; i=2: 8*2 + 160 = 176 > 170 -> cut [0,1],
#   restart at 2 with one element; i=3: 168; i=4: 176 -> cut [2,3];
#   i=5: 168; end -> [4,5].
# Intervals: p0 = [0,1], p1 = [2,3], p2 = [4,5].
#
# Superstep 0: every key (2, 4) lies outside its source's partition and E
# starts no rule: nothing. D <- input rows, old empty, in all three.
# Superstep 1:
#   pair (0,1): compute11(D0, D1): B(0,2) needs C at 2: D1 holds only D(2,4).
#     Nothing; tmp0, tmp1 empty.
#   pair (0,2): key 2 is outside p2; nothing. mergeAndDiff(p0): old0 = B02.
#   pair (1,2): compute11(D1, D2): D(2,4) x E(4,5) -> tmp1 C(2,5); spagpu
#     adds nothing more. mergeAndDiff(p1): old1 = D24, D1 = C25.
#   mergeAndDiff(p2): old2 = E45.
# Superstep 2:
#   pair (0,1): spagpu_b's first launch, compute11(old0, D1), is the
#     old_B x delta_C join B(0,2) x C(2,5) -> A(0,5); filter 1 skips it
#     (src/kernel.cu:556). D0 is empty, so no other launch has p0 keys, and
#     spagpu's old0 x tmp launch is skipped too (src/kernel.cu:644).
#   pair (0,2), (1,2): nothing (C starts no rule).
#   mergeAndDiff(p1): old1 = D24 C25, D1 empty -> every D empty, done.
# Result: the input plus C(2,5); A(0,5) is absent. NUM SUPERSTEPS: 2.
# The least fixpoint also holds A(0,5).
F3 = {
    "name": "cross_partition_old_x_delta_skipped",
    "n": 6,
    "num_labels": 5,
    "rules": [[1, 2, 3], [3, 4, 5]],
    "edges": [[0, 2, 2], [2, 4, 4], [4, 5, 5]],
    "heap_words": 680,
    "parts": [(0, 1), (2, 3), (4, 5)],
    "filter1": [[0, 2, 2], [2, 4, 4], [2, 5, 3], [4, 5, 5]],
    "filter1_supersteps": 2,
    "filter0": [[0, 2, 2], [0, 5, 1], [2, 4, 4], [2, 5, 3], [4, 5, 5]],
}

FIXTURES = [F1, F2, F3]
