// CPU oracle for the grammar-guided CFL-reachability closure.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker the GPU path is compared
// against and the CPU baseline timed beside it. Nothing in the product
// (graspan-g_amd/) links, loads or calls it.
//
// Semantics restated from the reference (Graspan-G, src/):
//   * rule table: int triples (lhs, rhs1, rhs2), 0 = absent, labels 1..L
//     (src/main.cpp:69-97; rule classes at src/kernel.cu:403-464):
//       rhs1 == 0            eps     A(v,v) for every v in [0,n)   (compute00, src/kernel.cu:445-464)
//       rhs1 != 0, rhs2 == 0 unary   A(u,w) <= B(u,w)              (compute10, src/kernel.cu:422-442)
//       both != 0            binary  A(u,w) <= B(u,v) & C(v,w)     (compute11/apply/decode/unionAll,
//                                                                   src/kernel.cu:297-362, 403-419)
//   * edges form a set (bitmap elements OR-ed together, src/kernel.cu:165-202, 234-295);
//   * with filter == 0 the result is the least fixpoint containing the input
//     edges (the schedule of src/main.cpp:764-906 and src/kernel.cu:466-780
//     joins every newly derived edge against every co-resident region, and the
//     pair sweep covers the rest), so a global semi-naive evaluation produces
//     the same set;
//   * end of a step: old <- old U delta, delta <- tmp - old (mergeAndDiff,
//     src/kernel.cu:943-985); done when every delta is empty
//     (src/main.cpp:896-905);
//   * FINAL NUM EDGES = sum of the per-vertex edge counts of old
//     (src/main.cpp:909-917, src/kernel.cu:782-839), kept here as uint64.
//
// "Mode G" below is global Jacobi semi-naive over sorted per-(label, vertex)
// rows:
//     cand(A,u) = U_{A::=B}   delta_B(u)
//               U_{A::=B C}  ( U_{v in delta_B(u)} old_C(v) U delta_C(v) )
//                            U ( U_{v in old_B(u)} delta_C(v) )
//     delta'(A,u) = cand(A,u) - old_A(u) - delta_A(u)
// It also records, per round, the algorithmic counters used for the HBM
// roofline figure:
//     P  probe reads: delta-key / old-key elements scanned plus every C-row
//        element read by the join
//     K  candidates emitted
//     O  elements of old_A(u) U delta_A(u) read for the difference, touched
//        (A,u) only
//     N  new unique edges.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <string>
#include <vector>

namespace {

struct Csr {
  std::vector<uint64_t> ptr;  // n + 1
  std::vector<uint32_t> col;  // sorted, unique within a row
  void empty(uint32_t n) {
    ptr.assign(size_t(n) + 1, 0);
    col.clear();
    col.shrink_to_fit();
  }
  uint64_t nnz() const { return col.size(); }
  uint64_t len(uint32_t u) const { return ptr[u + 1] - ptr[u]; }
  const uint32_t* row(uint32_t u) const { return col.data() + ptr[u]; }
};

struct Grammar {
  int L = 0;
  std::vector<int> eps;                                   // lhs (0-based)
  std::vector<std::vector<int>> unary;                    // [lhs] -> B
  std::vector<std::vector<std::pair<int, int>>> binary;   // [lhs] -> (B, C)
};

struct RoundStat {
  uint64_t P = 0, K = 0, O = 0, N = 0;
};

struct Result {
  uint32_t n = 0;
  int L = 0;
  std::vector<Csr> rel;
  std::vector<RoundStat> rounds;
  uint64_t input_unique = 0;  // |dedup input U eps loops|
  double seconds = 0.0;
  std::string err;
  // mode R only
  uint32_t supersteps = 0, num_parts = 0;
  std::vector<uint32_t> part_first, part_last;
};

// Build one CSR per label from COO; rows sorted + unique.
void build_csr(uint32_t n, int L, const std::vector<uint32_t>& src, const std::vector<uint32_t>& dst,
               const std::vector<uint8_t>& lab, std::vector<Csr>& out) {
  out.assign(L, Csr());
  const size_t m = src.size();
  for (int l = 0; l < L; ++l) out[l].ptr.assign(size_t(n) + 1, 0);
  for (size_t i = 0; i < m; ++i) out[lab[i]].ptr[src[i] + 1]++;
  for (int l = 0; l < L; ++l) {
    auto& p = out[l].ptr;
    for (uint32_t u = 0; u < n; ++u) p[u + 1] += p[u];
    out[l].col.resize(p[n]);
  }
  std::vector<std::vector<uint64_t>> fill(L);
  for (int l = 0; l < L; ++l) fill[l].assign(out[l].ptr.begin(), out[l].ptr.end() - 1);
  for (size_t i = 0; i < m; ++i) out[lab[i]].col[fill[lab[i]][src[i]]++] = dst[i];
  // sort + unique each row, then compact
  for (int l = 0; l < L; ++l) {
    Csr& c = out[l];
    std::vector<uint64_t> newlen(n);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t u = 0; u < int64_t(n); ++u) {
      uint32_t* b = c.col.data() + c.ptr[u];
      uint32_t* e = c.col.data() + c.ptr[u + 1];
      std::sort(b, e);
      newlen[u] = uint64_t(std::unique(b, e) - b);
    }
    std::vector<uint64_t> np(size_t(n) + 1, 0);
    for (uint32_t u = 0; u < n; ++u) np[u + 1] = np[u] + newlen[u];
    std::vector<uint32_t> nc(np[n]);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t u = 0; u < int64_t(n); ++u)
      std::memcpy(nc.data() + np[u], c.col.data() + c.ptr[u], newlen[u] * sizeof(uint32_t));
    c.ptr.swap(np);
    c.col.swap(nc);
  }
}

// old <- old U delta (disjoint sorted rows).
void merge_into(uint32_t n, Csr& old, const Csr& delta) {
  if (delta.nnz() == 0) return;
  std::vector<uint64_t> np(size_t(n) + 1, 0);
  for (uint32_t u = 0; u < n; ++u) np[u + 1] = np[u] + old.len(u) + delta.len(u);
  std::vector<uint32_t> nc(np[n]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < int64_t(n); ++u) {
    std::merge(old.row(u), old.row(u) + old.len(u), delta.row(u), delta.row(u) + delta.len(u),
               nc.data() + np[u]);
  }
  old.ptr.swap(np);
  old.col.swap(nc);
}

constexpr uint32_t kBlockRows = 256;

// One semi-naive round for every lhs label; next[A] receives delta'(A, .).
void round_g(const Grammar& g, uint32_t n, const std::vector<Csr>& old, const std::vector<Csr>& delta,
             std::vector<Csr>& next, RoundStat& st) {
  const int L = g.L;
  next.assign(L, Csr());
  const uint32_t nblocks = (n + kBlockRows - 1) / kBlockRows;
  uint64_t P = 0, K = 0, O = 0, N = 0;
  for (int A = 0; A < L; ++A) {
    Csr& out = next[A];
    out.ptr.assign(size_t(n) + 1, 0);
    const auto& un = g.unary[A];
    const auto& bi = g.binary[A];
    bool any = false;
    for (int B : un) any |= delta[B].nnz() > 0;
    for (auto& bc : bi) any |= delta[bc.first].nnz() > 0 || delta[bc.second].nnz() > 0;
    if (!any) continue;
    std::vector<std::vector<uint32_t>> blk(nblocks);
#pragma omp parallel reduction(+ : P, K, O, N)
    {
      std::vector<uint32_t> stamp(n, 0);
      uint32_t cur = 0;
      std::vector<uint32_t> cand;
#pragma omp for schedule(dynamic, 1)
      for (int64_t b = 0; b < int64_t(nblocks); ++b) {
        const uint32_t u0 = uint32_t(b) * kBlockRows;
        const uint32_t u1 = std::min<uint32_t>(n, u0 + kBlockRows);
        std::vector<uint32_t>& dst = blk[b];
        for (uint32_t u = u0; u < u1; ++u) {
          cand.clear();
          for (int B : un) {
            const Csr& dB = delta[B];
            cand.insert(cand.end(), dB.row(u), dB.row(u) + dB.len(u));
            P += dB.len(u);
          }
          for (auto& bc : bi) {
            const Csr& dB = delta[bc.first];
            const Csr& oC = old[bc.second];
            const Csr& dC = delta[bc.second];
            const uint32_t* kb = dB.row(u);
            const uint64_t kl = dB.len(u);
            P += kl;
            for (uint64_t i = 0; i < kl; ++i) {
              const uint32_t v = kb[i];
              cand.insert(cand.end(), oC.row(v), oC.row(v) + oC.len(v));
              cand.insert(cand.end(), dC.row(v), dC.row(v) + dC.len(v));
              P += oC.len(v) + dC.len(v);
            }
            if (dC.nnz() > 0) {
              const Csr& oB = old[bc.first];
              const uint32_t* ob = oB.row(u);
              const uint64_t ol = oB.len(u);
              P += ol;
              for (uint64_t i = 0; i < ol; ++i) {
                const uint32_t v = ob[i];
                cand.insert(cand.end(), dC.row(v), dC.row(v) + dC.len(v));
                P += dC.len(v);
              }
            }
          }
          if (cand.empty()) {
            next[A].ptr[u + 1] = 0;
            continue;
          }
          K += cand.size();
          if (++cur == 0) {  // stamp wrap-around
            std::fill(stamp.begin(), stamp.end(), 0);
            cur = 1;
          }
          const Csr& oA = old[A];
          const Csr& dA = delta[A];
          for (uint64_t i = 0; i < oA.len(u); ++i) stamp[oA.row(u)[i]] = cur;
          for (uint64_t i = 0; i < dA.len(u); ++i) stamp[dA.row(u)[i]] = cur;
          O += oA.len(u) + dA.len(u);
          const size_t before = dst.size();
          for (uint32_t w : cand) {
            if (stamp[w] != cur) {
              stamp[w] = cur;
              dst.push_back(w);
            }
          }
          std::sort(dst.begin() + before, dst.end());
          const uint64_t got = dst.size() - before;
          N += got;
          next[A].ptr[u + 1] = got;
        }
      }
    }
    auto& p = out.ptr;
    for (uint32_t u = 0; u < n; ++u) p[u + 1] += p[u];
    out.col.resize(p[n]);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < int64_t(nblocks); ++b) {
      if (!blk[b].empty())
        std::memcpy(out.col.data() + p[uint32_t(b) * kBlockRows], blk[b].data(), blk[b].size() * sizeof(uint32_t));
    }
  }
  for (int A = 0; A < L; ++A)
    if (next[A].ptr.empty()) next[A].empty(n);
  st.P = P;
  st.K = K;
  st.O = O;
  st.N = N;
}

bool parse_rules(int L, int R, const int32_t* rules, Grammar& g, std::string& err) {
  g = Grammar();
  g.L = L;
  g.unary.assign(L, {});
  g.binary.assign(L, {});
  for (int r = 0; r < R; ++r) {
    const int a = rules[3 * r], b = rules[3 * r + 1], c = rules[3 * r + 2];
    if (a < 1 || a > L || b < 0 || b > L || c < 0 || c > L || (b == 0 && c != 0)) {
      err = "bad rule " + std::to_string(r);
      return false;
    }
    if (b == 0)
      g.eps.push_back(a - 1);
    else if (c == 0)
      g.unary[a - 1].push_back(b - 1);
    else
      g.binary[a - 1].push_back({b - 1, c - 1});
  }
  return true;
}

Result* run_g(uint32_t n, int L, int R, const int32_t* rules, uint64_t m, const uint32_t* src,
              const uint32_t* dst, const uint8_t* lab, int max_rounds) {
  auto* res = new Result();
  res->n = n;
  res->L = L;
  Grammar g;
  if (!parse_rules(L, R, rules, g, res->err)) return res;
  for (uint64_t i = 0; i < m; ++i) {
    if (src[i] >= n || dst[i] >= n || lab[i] < 1 || lab[i] > L) {
      res->err = "bad edge " + std::to_string(i);
      return res;
    }
  }
  auto t0 = std::chrono::steady_clock::now();
  // delta_0 = dedup(input U eps loops)
  std::vector<uint32_t> s(src, src + m), d(dst, dst + m);
  std::vector<uint8_t> l(m);
  for (uint64_t i = 0; i < m; ++i) l[i] = uint8_t(lab[i] - 1);
  for (int A : g.eps) {
    for (uint32_t v = 0; v < n; ++v) {
      s.push_back(v);
      d.push_back(v);
      l.push_back(uint8_t(A));
    }
  }
  std::vector<Csr> old(L), delta, next;
  for (int A = 0; A < L; ++A) old[A].empty(n);
  build_csr(n, L, s, d, l, delta);
  std::vector<uint32_t>().swap(s);
  std::vector<uint32_t>().swap(d);
  std::vector<uint8_t>().swap(l);
  for (int A = 0; A < L; ++A) res->input_unique += delta[A].nnz();
  for (int it = 0; max_rounds <= 0 || it < max_rounds; ++it) {
    uint64_t dsz = 0;
    for (int A = 0; A < L; ++A) dsz += delta[A].nnz();
    if (dsz == 0) break;
    RoundStat st;
    round_g(g, n, old, delta, next, st);
    for (int A = 0; A < L; ++A) merge_into(n, old[A], delta[A]);
    delta.swap(next);
    res->rounds.push_back(st);
  }
  // a truncated run (max_rounds) folds the pending delta in as well
  for (int A = 0; A < L; ++A) merge_into(n, old[A], delta[A]);
  res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  res->rel.swap(old);
  return res;
}

// ---------------------------------------------------------------------------
// Mode R: the reference's own schedule, restated at the level of edge sets.
// Needed for filter != 0, whose result depends on the partitioning and the
// sweep order.
//
//   partitions   initializePartitions (src/main.cpp:141-175) over the degree
//                pre-pass of getGraphInfo (src/main.cpp:106-139): per source
//                the number of input lines and of distinct (dst/192, label)
//                "elements"; heap = DEVMEMUSED_MB 10000 -> 2,293,760,000
//                words (src/kernel.cu:35, src/kernel.h:26), partition limit =
//                heap/4 words of 8*elements + roundup32(vertices)*labels.
//   superstep 0  precomputation (src/main.cpp:764-804): per partition, repeat
//                spagpu_s (src/kernel.cu:466-521) + mergeAndDiff
//                (src/kernel.cu:943-985) until delta is empty, then
//                delta <- old, old <- {}.
//   superstep k  run_computation (src/main.cpp:806-858): for i, for j > i:
//                spagpu_b (src/kernel.cu:523-628) once, then spagpu
//                (src/kernel.cu:630-780) until nothing is added; then
//                mergeAndDiff(i). Done when every delta is empty
//                (src/main.cpp:892-906).
//   compute11(src, dst1, dst2, tmp): tmp_A(u) |= dstk_C(v) for every binary
//                rule A ::= B C, u in src's partition, v in src_B(u) and v in
//                dstk's partition (src/kernel.cu:297-362, 403-419).
//   compute10    tmp_A(u) |= from_B(u) for unary rules (src/kernel.cu:422-442).
//   compute00    tmp_A(u) |= {u} for eps rules (src/kernel.cu:445-464).
// A region that does not exist (size 0 in the reference) joins nothing; the
// existence flags below follow oldSize/deltaSize/tmpSize.
namespace moder {

using Rows = std::vector<std::vector<uint32_t>>;  // [label * n + u], sorted unique

struct Part {
  uint32_t first = 0, last = 0;
  bool has_old = false, has_dl = false, has_tmp = false;
};

struct State {
  uint32_t n = 0;
  int L = 0;
  const Grammar* g = nullptr;
  int filter = 0;
  Rows old, dl, tmp;
  std::vector<Part> parts;
  bool changed = false;
  uint32_t supersteps = 0;
};

enum Region { kOld, kDl, kTmp };

Rows& region(State& s, Region r) { return r == kOld ? s.old : (r == kDl ? s.dl : s.tmp); }
bool exists(const Part& p, Region r) { return r == kOld ? p.has_old : (r == kDl ? p.has_dl : p.has_tmp); }

// tmp_A(u) |= add (sorted unique); sets s.changed when something is new.
void add_to(State& s, size_t row, std::vector<uint32_t>& add) {
  if (add.empty()) return;
  std::sort(add.begin(), add.end());
  add.erase(std::unique(add.begin(), add.end()), add.end());
  std::vector<uint32_t>& t = s.tmp[row];
  std::vector<uint32_t> out;
  out.reserve(t.size() + add.size());
  std::set_union(t.begin(), t.end(), add.begin(), add.end(), std::back_inserter(out));
  if (out.size() != t.size()) s.changed = true;
  t.swap(out);
}

struct Dst {
  Region r;
  int part;
};

// compute11(src region of partition p, dsts..., tmp of p). All reads see the
// regions as they were when the operator started (the result of the
// enclosing loops does not depend on it: every loop runs to its fixpoint or
// applies operators that read no tmp).
void join(State& s, int p, Region src, std::initializer_list<Dst> dsts_in) {
  const Part& P = s.parts[p];
  if (!exists(P, src)) return;
  const uint32_t n = s.n;
  Rows& S = region(s, src);
  const std::vector<Dst> dsts(dsts_in);
  for (int A = 0; A < s.L; ++A) {
    if (s.g->binary[A].empty()) continue;
    // candidates of every row of partition p first (reads only), then the
    // row-wise unions (each row written by one thread)
    const uint32_t cnt = P.last - P.first + 1;
    std::vector<std::vector<uint32_t>> add(cnt);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t k = 0; k < int64_t(cnt); ++k) {
      const uint32_t u = P.first + uint32_t(k);
      for (auto& bc : s.g->binary[A]) {
        for (uint32_t v : S[size_t(bc.first) * n + u]) {
          for (const Dst& d : dsts) {
            const Part& Q = s.parts[d.part];
            if (!exists(Q, d.r) || v < Q.first || v > Q.last) continue;
            const auto& lst = region(s, d.r)[size_t(bc.second) * n + v];
            add[k].insert(add[k].end(), lst.begin(), lst.end());
          }
        }
      }
    }
    bool ch = false;
#pragma omp parallel for schedule(dynamic, 64) reduction(|| : ch)
    for (int64_t k = 0; k < int64_t(cnt); ++k) {
      if (add[k].empty()) continue;
      std::vector<uint32_t>& a = add[k];
      std::sort(a.begin(), a.end());
      a.erase(std::unique(a.begin(), a.end()), a.end());
      std::vector<uint32_t>& t = s.tmp[size_t(A) * n + P.first + uint32_t(k)];
      std::vector<uint32_t> out;
      out.reserve(t.size() + a.size());
      std::set_union(t.begin(), t.end(), a.begin(), a.end(), std::back_inserter(out));
      if (out.size() != t.size()) ch = true;
      t.swap(out);
    }
    if (ch) s.changed = true;
  }
}

// compute10 from region `from` of partition p.
void unary(State& s, int p, Region from) {
  const Part& P = s.parts[p];
  if (!exists(P, from)) return;
  const uint32_t n = s.n;
  std::vector<std::pair<size_t, std::vector<uint32_t>>> adds;
  for (int A = 0; A < s.L; ++A)
    for (int B : s.g->unary[A])
      for (uint32_t u = P.first; u <= P.last; ++u) {
        const auto& lst = region(s, from)[size_t(B) * n + u];
        if (!lst.empty()) adds.emplace_back(size_t(A) * n + u, lst);
      }
  for (auto& a : adds) add_to(s, a.first, a.second);
}

void eps(State& s, int p) {
  const Part& P = s.parts[p];
  for (int A : s.g->eps)
    for (uint32_t u = P.first; u <= P.last; ++u) {
      std::vector<uint32_t> one{u};
      add_to(s, size_t(A) * s.n + u, one);
    }
}

// mergeAndDiff (src/kernel.cu:943-985) for partition p.
void merge_and_diff(State& s, int p) {
  Part& P = s.parts[p];
  const uint32_t n = s.n;
  if (P.has_dl) {
    for (int A = 0; A < s.L; ++A)
      for (uint32_t u = P.first; u <= P.last; ++u) {
        const size_t r = size_t(A) * n + u;
        if (!P.has_old) {
          s.old[r].swap(s.dl[r]);
        } else {
          std::vector<uint32_t> out;
          std::set_union(s.old[r].begin(), s.old[r].end(), s.dl[r].begin(), s.dl[r].end(), std::back_inserter(out));
          s.old[r].swap(out);
        }
        s.dl[r].clear();
      }
    P.has_old = true;
    P.has_dl = false;
  }
  if (P.has_tmp) {
    bool any = false;
    for (int A = 0; A < s.L; ++A)
      for (uint32_t u = P.first; u <= P.last; ++u) {
        const size_t r = size_t(A) * n + u;
        std::vector<uint32_t> out;
        std::set_difference(s.tmp[r].begin(), s.tmp[r].end(), s.old[r].begin(), s.old[r].end(),
                            std::back_inserter(out));
        any |= !out.empty();
        s.dl[r].swap(out);
        s.tmp[r].clear();
      }
    P.has_dl = any;
    P.has_tmp = false;
  }
}

// Planted mistakes, for the tests that show the hand-derived filter
// fixtures can see them (ORC_MODE_R_MUTATION; 0 = the transcription as is):
//   1  with filter != 0, skip delta_B x old_C and run old_B x delta_C instead
//   2  drop spagpu's tmp x tmp launches
//   3  run spagpu's loop once instead of until nothing changes
int mutation() {
  const char* m = std::getenv("ORC_MODE_R_MUTATION");
  return m ? std::atoi(m) : 0;
}

void spagpu_s(State& s, int p) {
  Part& P = s.parts[p];
  const bool swap = s.filter != 0 && mutation() == 1;
  P.has_tmp = true;
  if (!P.has_old) eps(s, p);
  if (s.filter == 0) unary(s, p, kDl);
  if (P.has_old) {
    if (s.filter == 0 || swap) join(s, p, kOld, {{kDl, p}});
    if (swap)
      join(s, p, kDl, {{kDl, p}});
    else
      join(s, p, kDl, {{kOld, p}, {kDl, p}});
  } else {
    join(s, p, kDl, {{kDl, p}});
  }
}

void spagpu_b(State& s, int i, int j) {
  const bool swap = s.filter != 0 && mutation() == 1;
  s.parts[i].has_tmp = true;
  s.parts[j].has_tmp = true;
  if (s.filter == 0 || swap) join(s, i, kOld, {{kDl, j}});
  if (swap)
    join(s, i, kDl, {{kDl, j}});
  else
    join(s, i, kDl, {{kOld, j}, {kDl, j}});
  if (s.filter == 0 || swap) join(s, j, kOld, {{kDl, i}});
  if (swap)
    join(s, j, kDl, {{kDl, i}});
  else
    join(s, j, kDl, {{kOld, i}, {kDl, i}});
}

void spagpu(State& s, int i, int j) {
  const int mut = mutation();
  for (;;) {
    s.changed = false;
    for (int side = 0; side < 2; ++side) {
      const int p = side ? j : i, q = side ? i : j;
      if (s.filter == 0) join(s, p, kOld, {{kTmp, i}, {kTmp, j}});
      join(s, p, kDl, {{kTmp, i}, {kTmp, j}});
      join(s, p, kTmp, {{kOld, p}, {kDl, p}});
      join(s, p, kTmp, {{kOld, q}, {kDl, q}});
      if (s.filter == 0) unary(s, p, kTmp);
      if (mut != 2) join(s, p, kTmp, {{kTmp, i}, {kTmp, j}});
    }
    if (!s.changed || mut == 3) break;
  }
}

// initializePartitions over the getGraphInfo degree pre-pass. Returns false
// for the degenerate splits where the reference's unsigned arithmetic breaks
// (a first vertex holding more than half the elements, or no elements).
bool make_partitions(State& s, uint64_t m, const uint32_t* src, const uint32_t* dst, const uint8_t* lab,
                     std::string& err) {
  const uint32_t n = s.n;
  std::vector<uint64_t> edges(n, 0), elems(n, 0);
  {
    std::vector<std::vector<uint64_t>> keys(n);
    for (uint64_t i = 0; i < m; ++i) {
      edges[src[i]]++;
      keys[src[i]].push_back((uint64_t(dst[i] / 192) << 8) | lab[i]);
    }
    for (uint32_t u = 0; u < n; ++u) {
      auto& k = keys[u];
      std::sort(k.begin(), k.end());
      elems[u] = uint64_t(std::unique(k.begin(), k.end()) - k.begin());
    }
  }
  // the heap of allocateElementPool (src/kernel.cu:35) in 32-bit words;
  // GG_REF_HEAP_WORDS replaces it so that tests reach the multi-partition cut
  uint64_t heap = (10000ull / 8 * 7) * 1024 * 256;
  if (const char* h = std::getenv("GG_REF_HEAP_WORDS"))
    if (std::strtoull(h, nullptr, 10)) heap = std::strtoull(h, nullptr, 10);
  const uint64_t limit = heap / 4;
  uint64_t t1 = 0, t3 = 0;
  auto rup32 = [](uint64_t x) { return (x + 31) / 32 * 32; };
  for (uint32_t i = 0; i < n; ++i) {
    t1 += elems[i];
    if (t1 * 8 + rup32(i - t3 + 1) * uint64_t(s.L) > limit) {
      if (i == 0 || uint64_t(i - 1) < t3) {
        err = "degenerate partitioning";
        return false;
      }
      Part p;
      p.first = uint32_t(t3);
      p.last = i - 1;
      s.parts.push_back(p);
      t1 = elems[i];
      t3 = i;
    }
  }
  Part last;
  last.first = uint32_t(t3);
  last.last = n - 1;
  s.parts.push_back(last);
  if (s.parts.size() == 1) {
    uint64_t t4 = 0;
    bool split = false;
    for (uint32_t i = 0; i < n; ++i) {
      t4 += elems[i];
      if (t4 > t1 / 2) {
        if (i == 0) break;
        s.parts[0].last = i - 1;
        Part p2;
        p2.first = i;
        p2.last = n - 1;
        s.parts.push_back(p2);
        split = true;
        break;
      }
    }
    if (!split) {
      err = "degenerate partitioning (the reference's small-graph split needs a second vertex interval)";
      return false;
    }
  }
  return true;
}

}  // namespace moder

Result* run_r(uint32_t n, int L, int R, const int32_t* rules, uint64_t m, const uint32_t* src, const uint32_t* dst,
              const uint8_t* lab, int filter, bool parts_only = false) {
  auto* res = new Result();
  res->n = n;
  res->L = L;
  Grammar g;
  if (!parse_rules(L, R, rules, g, res->err)) return res;
  for (uint64_t i = 0; i < m; ++i)
    if (src[i] >= n || dst[i] >= n || lab[i] < 1 || lab[i] > L) {
      res->err = "bad edge " + std::to_string(i);
      return res;
    }
  auto t0 = std::chrono::steady_clock::now();
  moder::State s;
  s.n = n;
  s.L = L;
  s.g = &g;
  s.filter = filter;
  if (!moder::make_partitions(s, m, src, dst, lab, res->err)) return res;
  if (parts_only) {
    res->num_parts = uint32_t(s.parts.size());
    for (auto& p : s.parts) {
      res->part_first.push_back(p.first);
      res->part_last.push_back(p.last);
    }
    res->rel.assign(L, Csr());
    for (auto& c : res->rel) c.ptr.assign(size_t(n) + 1, 0);
    return res;
  }
  const size_t rows = size_t(L) * n;
  s.old.assign(rows, {});
  s.dl.assign(rows, {});
  s.tmp.assign(rows, {});
  for (uint64_t i = 0; i < m; ++i) s.dl[size_t(lab[i] - 1) * n + src[i]].push_back(dst[i]);
  for (auto& r : s.dl) {
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
  }
  for (auto& p : s.parts) p.has_dl = true;  // loaded from the graph file
  // superstep 0
  for (size_t i = 0; i < s.parts.size(); ++i) {
    for (;;) {
      moder::spagpu_s(s, int(i));
      moder::merge_and_diff(s, int(i));
      if (!s.parts[i].has_dl) {
        moder::Part& P = s.parts[i];
        for (int A = 0; A < L; ++A)
          for (uint32_t u = P.first; u <= P.last; ++u) s.dl[size_t(A) * n + u].swap(s.old[size_t(A) * n + u]);
        P.has_dl = P.has_old;
        P.has_old = false;
        break;
      }
    }
  }
  // supersteps 1..
  for (;;) {
    ++s.supersteps;
    for (size_t i = 0; i < s.parts.size(); ++i) {
      for (size_t j = i + 1; j < s.parts.size(); ++j) {
        moder::spagpu_b(s, int(i), int(j));
        moder::spagpu(s, int(i), int(j));
      }
      moder::merge_and_diff(s, int(i));
    }
    bool done = true;
    for (auto& p : s.parts) done &= !p.has_dl;
    if (done) break;
  }
  res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  // closure = every partition's old region
  res->rel.assign(L, Csr());
  for (int A = 0; A < L; ++A) {
    Csr& c = res->rel[A];
    c.ptr.assign(size_t(n) + 1, 0);
    for (uint32_t u = 0; u < n; ++u) c.ptr[u + 1] = c.ptr[u] + s.old[size_t(A) * n + u].size();
    c.col.resize(c.ptr[n]);
    for (uint32_t u = 0; u < n; ++u)
      std::copy(s.old[size_t(A) * n + u].begin(), s.old[size_t(A) * n + u].end(), c.col.begin() + long(c.ptr[u]));
  }
  res->supersteps = s.supersteps;
  res->num_parts = uint32_t(s.parts.size());
  for (auto& p : s.parts) {
    res->part_first.push_back(p.first);
    res->part_last.push_back(p.last);
  }
  return res;
}

}  // namespace

extern "C" {

// Runs the closure. filter must be 0 (mode G). max_rounds <= 0 runs to the
// fixpoint; a positive value stops early (used only for bounded timing
// samples). threads <= 0 keeps the OpenMP default.
void* orc_run(uint32_t n, int L, int R, const int32_t* rules, uint64_t m, const uint32_t* src, const uint32_t* dst,
              const uint8_t* lab, int filter, int threads, int max_rounds) {
  if (threads > 0) omp_set_num_threads(threads);
  if (filter != 0) {
    auto* r = new Result();
    r->err = "oracle mode G handles filter == 0 only";
    return r;
  }
  return run_g(n, L, R, rules, m, src, dst, lab, max_rounds);
}

// Mode R (the reference's partitioned schedule); any filter value.
void* orc_run_r(uint32_t n, int L, int R, const int32_t* rules, uint64_t m, const uint32_t* src, const uint32_t* dst,
                const uint8_t* lab, int filter) {
  return run_r(n, L, R, rules, m, src, dst, lab, filter);
}

// Mode R's vertex intervals only (no closure is computed).
void* orc_parts_r(uint32_t n, int L, int R, const int32_t* rules, uint64_t m, const uint32_t* src, const uint32_t* dst,
                  const uint8_t* lab) {
  return run_r(n, L, R, rules, m, src, dst, lab, 1, true);
}

// Mode R schedule facts: supersteps, partition count, partition bounds.
uint32_t orc_supersteps(void* h) { return static_cast<Result*>(h)->supersteps; }
uint32_t orc_num_parts(void* h) { return static_cast<Result*>(h)->num_parts; }
void orc_parts(void* h, uint32_t* first, uint32_t* last) {
  auto* r = static_cast<Result*>(h);
  for (size_t i = 0; i < r->part_first.size(); ++i) {
    first[i] = r->part_first[i];
    last[i] = r->part_last[i];
  }
}

const char* orc_error(void* h) { return static_cast<Result*>(h)->err.c_str(); }

uint64_t orc_total(void* h) {
  auto* r = static_cast<Result*>(h);
  uint64_t t = 0;
  for (auto& c : r->rel) t += c.nnz();
  return t;
}

uint64_t orc_input_unique(void* h) { return static_cast<Result*>(h)->input_unique; }

void orc_counts(void* h, uint64_t* per_label) {
  auto* r = static_cast<Result*>(h);
  for (int l = 0; l < r->L; ++l) per_label[l] = r->rel[l].nnz();
}

int orc_num_rounds(void* h) { return int(static_cast<Result*>(h)->rounds.size()); }

void orc_round_stats(void* h, uint64_t* out) {
  auto* r = static_cast<Result*>(h);
  for (size_t i = 0; i < r->rounds.size(); ++i) {
    out[4 * i + 0] = r->rounds[i].P;
    out[4 * i + 1] = r->rounds[i].K;
    out[4 * i + 2] = r->rounds[i].O;
    out[4 * i + 3] = r->rounds[i].N;
  }
}

double orc_seconds(void* h) { return static_cast<Result*>(h)->seconds; }

// Per-vertex total out-degree over all labels (the reference's degree[old]
// column, src/kernel.cu:818-839), uint64.
void orc_degrees(void* h, uint64_t* deg) {
  auto* r = static_cast<Result*>(h);
  for (uint32_t u = 0; u < r->n; ++u) {
    uint64_t s = 0;
    for (auto& c : r->rel) s += c.len(u);
    deg[u] = s;
  }
}

// Writes the closure as (src, dst, label) uint32 triples sorted
// lexicographically, label ids 1-based. Returns the triple count; writes
// nothing if cap is too small.
uint64_t orc_export(void* h, uint32_t* buf, uint64_t cap) {
  auto* r = static_cast<Result*>(h);
  const uint32_t n = r->n;
  const uint64_t total = orc_total(h);
  if (buf == nullptr || cap < total) return total;
  std::vector<uint64_t> off(size_t(n) + 1, 0);
  for (uint32_t u = 0; u < n; ++u) {
    uint64_t s = 0;
    for (auto& c : r->rel) s += c.len(u);
    off[u + 1] = off[u] + s;
  }
#pragma omp parallel
  {
    std::vector<uint64_t> tmp;
#pragma omp for schedule(dynamic, 256)
    for (int64_t u = 0; u < int64_t(n); ++u) {
      tmp.clear();
      for (int l = 0; l < r->L; ++l) {
        const Csr& c = r->rel[l];
        for (uint64_t i = 0; i < c.len(uint32_t(u)); ++i) tmp.push_back((uint64_t(c.row(uint32_t(u))[i]) << 8) | uint64_t(l + 1));
      }
      std::sort(tmp.begin(), tmp.end());
      uint32_t* o = buf + 3 * off[u];
      for (size_t i = 0; i < tmp.size(); ++i) {
        o[3 * i] = uint32_t(u);
        o[3 * i + 1] = uint32_t(tmp[i] >> 8);
        o[3 * i + 2] = uint32_t(tmp[i] & 0xFF);
      }
    }
  }
  return total;
}

// Order-independent checksum: sum mod 2^64 of h(src, dst, label) over the
// closure, h = splitmix64 finaliser of a linear mix of the ids (label 1-based).
static inline uint64_t orc_hash(uint32_t s, uint32_t d, uint32_t l) {
  uint64_t x = uint64_t(s) * 0x9E3779B97F4A7C15ull ^ uint64_t(d) * 0xC2B2AE3D27D4EB4Full ^
               uint64_t(l) * 0x165667B19E3779F9ull;
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return x;
}

uint64_t orc_checksum(void* h) {
  auto* r = static_cast<Result*>(h);
  uint64_t total = 0;
  for (int l = 0; l < r->L; ++l) {
    const Csr& c = r->rel[l];
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : total)
    for (int64_t u = 0; u < int64_t(r->n); ++u)
      for (uint64_t i = 0; i < c.len(uint32_t(u)); ++i) total += orc_hash(uint32_t(u), c.row(uint32_t(u))[i], uint32_t(l + 1));
  }
  return total;
}

void orc_free(void* h) { delete static_cast<Result*>(h); }

}  // extern "C"
