// Batched closure of many independent graphs (gg_batch_*; no counterpart in
// the reference, which closes one graph per process).
//
// The graphs of one launch group are laid side by side: graph g's vertices
// are shifted by v0[g], and each distinct grammar (same rule triples and label
// count) gets its own range of combined label ids. Eps rules are replaced by
// explicit self-loop edges on the vertices of the graphs that use them, so
// the combined grammar has only unary and binary rules and no rule can touch
// another graph's vertices or labels. The closure of this disjoint union is
// the disjoint union of the per-graph closures, computed by one engine run
// (one launch sequence per semi-naive round for the whole group).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "engine.h"
#include "formats.h"
#include "graspan_g.h"

struct gg_ctx;
namespace gg {
Engine* ctx_engine(gg_ctx* c, int& status, std::string*& err);
uint64_t ctx_memory(gg_ctx* c);
}

struct gg_batch {
  gg_ctx* ctx = nullptr;
  struct G {
    std::vector<int32_t> rules;
    int L = 0;
    uint32_t n = 0;
    std::vector<uint32_t> src, dst;
    std::vector<uint8_t> lab;
    // results
    std::vector<uint64_t> counts;
    uint64_t total = 0, checksum = 0;
    std::vector<uint32_t> triples;
    bool done = false;
  };
  std::vector<G> graphs;
  int groups = 0;
  std::string err;
};

namespace {

int fail(gg_batch* b, int st, const std::string& m) {
  b->err = m;
  return st;
}

void add_stats(gg_stats& a, const gg_stats& b) {
  a.rounds = std::max(a.rounds, b.rounds);
  a.retries += b.retries;
  a.input_unique += b.input_unique;
  a.total_edges += b.total_edges;
  a.fixpoint_ms += b.fixpoint_ms;
  a.expand_ms += b.expand_ms;
  a.compact_ms += b.compact_ms;
  a.merge_ms += b.merge_ms;
  a.other_ms += b.other_ms;
  a.expand_launches += b.expand_launches;
  a.probe_reads += b.probe_reads;
  a.candidates += b.candidates;
  a.old_reads += b.old_reads;
  a.new_edges += b.new_edges;
  a.device_bytes_peak = std::max(a.device_bytes_peak, b.device_bytes_peak);
  a.pair_reruns += b.pair_reruns;
  a.light_rows += b.light_rows;
}

int batch_run(gg_batch* b, uint64_t max_vertices, int keep, gg_stats* stats, std::string*& cerr) {
  int st = GG_OK;
  gg::Engine* e = gg::ctx_engine(b->ctx, st, cerr);
  if (!e) return fail(b, st, cerr ? *cerr : "no device");
  gg_stats total{};
  b->groups = 0;
  const size_t G = b->graphs.size();
  // Graphs are ordered by grammar before they are cut into launch groups: a
  // group's rows are its vertices times all its labels, so a group that
  // mixes fewer grammars carries fewer empty rows through every round.
  std::vector<size_t> ord(G);
  for (size_t k = 0; k < G; ++k) ord[k] = k;
  std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) {
    const auto& gx = b->graphs[x];
    const auto& gy = b->graphs[y];
    return std::make_pair(gx.L, gx.rules) < std::make_pair(gy.L, gy.rules);
  });
  // A group that runs out of device memory is cut in half and run again
  // (its closure size was not known in advance); later groups keep the
  // smaller graph count.
  size_t cap = G;
  // With max_vertices == 0 a group's vertex count is planned from device
  // memory: a group may use a tuned share of it (0.4 by default), at the peak device bytes
  // per vertex measured on the last group of the same grammar (a grammar not
  // yet seen starts with kFirstGroupVertices). Groups that run close to the
  // device's limit spend their time re-allocating instead of computing.
  const double kGroupMemShare = gg::tuned_group_mem_share();
  constexpr uint64_t kFirstGroupVertices = 2u << 20;
  const uint64_t mem = max_vertices ? 0 : gg::ctx_memory(b->ctx);
  std::map<std::pair<int, std::vector<int32_t>>, double> bytes_per_vertex;
  auto vertex_cap = [&](const std::pair<int, std::vector<int32_t>>& key) -> uint64_t {
    if (max_vertices) return max_vertices;
    auto it = bytes_per_vertex.find(key);
    if (it == bytes_per_vertex.end() || !mem) return kFirstGroupVertices;
    return std::max<uint64_t>(1, uint64_t(double(mem) * kGroupMemShare / std::max(it->second, 1.0)));
  };
  // GG_BATCH_TRACE=1: one stderr line per launch group (sizes and the host
  // time of each phase)
  const char* tr = std::getenv("GG_BATCH_TRACE");
  const bool trace = tr && tr[0] == '1';
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto ms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  size_t i = 0;
  while (i < G) {
    const auto t_start = now();
    // one launch group: graphs ord[i..j-1]
    // A group's combined grammar must stay within the engine's limits: at
    // most kMaxLabels labels, kMaxRulesPerKind unary and binary rules (eps
    // rules become loops), and labels * vertices and pair slots * vertices
    // below 2^32. A grammar that would pass one of them starts a new group.
    std::map<std::pair<int, std::vector<int32_t>>, uint32_t> grammar_base;
    uint32_t labels = 0;
    int unary = 0, binary = 0, slots = 0;
    uint64_t verts = 0;
    uint64_t vcap = ~0ull;  // the smallest vertex cap of the group's grammars
    size_t j = i;
    for (; j < G && j - i < cap; ++j) {
      const auto& g = b->graphs[ord[j]];
      auto key = std::make_pair(g.L, g.rules);
      const bool fresh = grammar_base.find(key) == grammar_base.end();
      gg::RuleKinds k{};
      if (fresh) k = gg::rule_kinds(g.rules.data(), int(g.rules.size() / 3));
      const uint32_t nl = labels + (fresh ? uint32_t(g.L) : 0u);
      const int nu = unary + k.unary, nb = binary + k.binary, ns = slots + k.slots;
      const uint64_t nv = verts + g.n;
      const uint64_t lim = nl ? (0xFFFFFFFEull / nl) : nv;
      const uint64_t slim = ns ? (0xFFFFFFFEull / uint64_t(ns)) : nv;
      const uint64_t ncap = std::min(vcap, vertex_cap(key));
      const bool hard = nl > uint32_t(gg::kMaxLabels) || nu > gg::kMaxRulesPerKind || nb > gg::kMaxRulesPerKind ||
                        nv > lim || nv > slim;
      if ((hard || nv > ncap) && j > i) break;  // a graph over the vertex cap alone still runs alone
      if (hard) return fail(b, GG_E_UNSUPPORTED, "graph " + std::to_string(ord[j]) + " alone exceeds the launch limits");
      if (fresh) {
        vcap = ncap;
        grammar_base[key] = labels;
        labels = nl;
        unary = nu;
        binary = nb;
        slots = ns;
      }
      verts = nv;
    }
    // combined grammar
    std::vector<int32_t> rules;
    for (auto& kv : grammar_base) {
      const auto& r = kv.first.second;
      for (size_t k = 0; k + 2 < r.size(); k += 3) {
        if (r[k + 1] == 0) continue;  // eps: explicit loops below
        rules.push_back(r[k] + int32_t(kv.second));
        rules.push_back(r[k + 1] + int32_t(kv.second));
        rules.push_back(r[k + 2] ? r[k + 2] + int32_t(kv.second) : 0);
      }
    }
    gg::Grammar cg;
    int s = gg::grammar_from_rules(int(rules.size() / 3), rules.data(), int(std::max(labels, 1u)), cg, b->err);
    if (s != GG_OK) return s;
    // super graph: every graph's edges shifted to its vertex range, plus the
    // eps loops its grammar asks for (sized first, then filled in place)
    gg::Graph sg;
    sg.n = uint32_t(verts);
    uint64_t total_m = 0;
    for (size_t k = i; k < j; ++k) {
      const auto& g = b->graphs[ord[k]];
      total_m += g.src.size();
      for (size_t r = 0; r + 2 < g.rules.size(); r += 3)
        if (g.rules[r + 1] == 0) total_m += g.n;
    }
    sg.src.resize(total_m);
    sg.dst.resize(total_m);
    sg.lab.resize(total_m);
    // each graph's slice of the union: its vertex base, label base and first
    // position, then the slices are filled by a few host threads
    std::vector<uint32_t> v0(1, 0), l0, lg;
    std::vector<uint64_t> at0;
    uint64_t at = 0;
    for (size_t k = i; k < j; ++k) {
      const auto& g = b->graphs[ord[k]];
      const uint32_t lb = grammar_base[std::make_pair(g.L, g.rules)];
      at0.push_back(at);
      at += g.src.size();
      for (size_t r = 0; r + 2 < g.rules.size(); r += 3)
        if (g.rules[r + 1] == 0) at += g.n;
      l0.push_back(lb);
      lg.push_back(uint32_t(g.L));
      v0.push_back(v0.back() + g.n);
    }
    auto fill = [&](size_t k0, size_t k1) {
      for (size_t k = k0; k < k1; ++k) {
        const auto& g = b->graphs[ord[i + k]];
        const uint32_t base = v0[k], lb = l0[k];
        uint64_t p = at0[k];
        const size_t m = g.src.size();
        uint32_t* ps = sg.src.data() + p;
        uint32_t* pd = sg.dst.data() + p;
        uint8_t* pl = sg.lab.data() + p;
        for (size_t e2 = 0; e2 < m; ++e2) {
          ps[e2] = g.src[e2] + base;
          pd[e2] = g.dst[e2] + base;
          pl[e2] = uint8_t(g.lab[e2] + lb);
        }
        p += m;
        for (size_t r = 0; r + 2 < g.rules.size(); r += 3) {
          if (g.rules[r + 1] != 0) continue;
          const uint8_t el = uint8_t(g.rules[r] + int32_t(lb));
          for (uint32_t v = 0; v < g.n; ++v) {
            sg.src[p + v] = base + v;
            sg.dst[p + v] = base + v;
            sg.lab[p + v] = el;
          }
          p += g.n;
        }
      }
    };
    const size_t ng = j - i;
    const size_t nt = std::min<size_t>({ng, 16, std::max(1u, std::thread::hardware_concurrency())});
    if (nt <= 1 || total_m < (1u << 20)) {
      fill(0, ng);
    } else {
      std::vector<std::thread> pool;
      for (size_t t = 0; t < nt; ++t) pool.emplace_back(fill, ng * t / nt, ng * (t + 1) / nt);
      for (auto& th : pool) th.join();
    }
    const auto t_built = now();
    if ((s = gg::engine_set_grammar(e, cg, b->err)) != GG_OK) return s;
    const uint64_t group_m = sg.src.size();
    s = gg::engine_set_graph(e, std::move(sg), b->err);
    const auto t_up = now();
    gg_stats one{};
    if (s == GG_OK) s = gg::engine_run(e, 0, &one, b->err);
    const auto t_run = now();
    if (s == GG_E_NOMEM && j - i > 1) {
      if (trace)
        std::fprintf(stderr, "[gg-batch] group of %zu graphs, %llu vertices: out of device memory after %.0f ms; "
                     "splitting\n", j - i, (unsigned long long)verts, ms(t_start, t_run));
      gg::engine_release(e);
      cap = (j - i) / 2;
      continue;
    }
    if (s != GG_OK) return s;
    add_stats(total, one);
    for (auto& kv : grammar_base) bytes_per_vertex[kv.first] = double(one.device_bytes_peak) / double(verts);
    std::vector<uint64_t> counts, sums;
    if ((s = gg::engine_slices(e, v0, l0, lg, counts, sums, b->err)) != GG_OK) return s;
    if (trace)
      std::fprintf(stderr,
                   "[gg-batch] group %d: %zu graphs, %llu vertices, %u labels, %llu edges -> %llu closure edges, "
                   "%u rounds; build %.0f ms, upload %.0f ms, run %.0f ms (fixpoint %.0f: expand %.0f merge %.0f "
                   "compact %.0f other %.0f), results %.0f ms, device peak %.1f GB\n",
                   b->groups, j - i, (unsigned long long)verts, labels, (unsigned long long)group_m,
                   (unsigned long long)one.total_edges, one.rounds, ms(t_start, t_built), ms(t_built, t_up),
                   ms(t_up, t_run), one.fixpoint_ms, one.expand_ms, one.merge_ms, one.compact_ms, one.other_ms,
                   ms(t_run, now()), double(one.device_bytes_peak) / 1e9);
    size_t c = 0;
    for (size_t k = i; k < j; ++k) {
      auto& g = b->graphs[ord[k]];
      g.counts.assign(counts.begin() + long(c), counts.begin() + long(c + size_t(g.L)));
      c += size_t(g.L);
      g.total = 0;
      for (uint64_t x : g.counts) g.total += x;
      g.checksum = sums[k - i];
      g.triples.clear();
      if (keep) {
        g.triples.resize(size_t(g.total) * 3);
        uint64_t cnt = 0;
        if ((s = gg::engine_export_slice(e, v0[k - i], v0[k - i + 1], l0[k - i], lg[k - i], g.triples.data(), g.total,
                                         &cnt, b->err)) != GG_OK)
          return s;
      }
      g.done = true;
    }
    ++b->groups;
    i = j;
  }
  if (stats) *stats = total;
  return GG_OK;
}

}  // namespace

extern "C" {

int gg_batch_create(gg_ctx* ctx, gg_batch** out) {
  if (!ctx || !out) return GG_E_ARG;
  auto* b = new gg_batch();
  b->ctx = ctx;
  *out = b;
  return GG_OK;
}

int gg_batch_destroy(gg_batch* b) {
  if (!b) return GG_E_ARG;
  delete b;
  return GG_OK;
}

int gg_batch_add(gg_batch* b, int num_rules, const int32_t* rules, int num_labels, uint32_t n, uint64_t m,
                 const uint32_t* src, const uint32_t* dst, const uint8_t* lab, int* graph_id) {
  if (!b) return GG_E_ARG;
  gg::Grammar tmp;
  int s = gg::grammar_from_rules(num_rules, rules, num_labels, tmp, b->err);
  if (s != GG_OK) return s;
  if (n == 0) return fail(b, GG_E_ARG, "graph has no vertices");
  if (m && (!src || !dst || !lab)) return fail(b, GG_E_ARG, "null edge array");
  for (uint64_t i = 0; i < m; ++i) {
    if (src[i] >= n || dst[i] >= n) return fail(b, GG_E_ARG, "edge " + std::to_string(i) + ": vertex out of range");
    if (lab[i] < 1 || lab[i] > num_labels) return fail(b, GG_E_ARG, "edge " + std::to_string(i) + ": bad label");
  }
  gg_batch::G g;
  // the graph is closed to its least fixpoint: its rules as planned for it
  // (eps rules stay, and become loop edges when the group is built)
  g.rules = gg::plan_fixpoint_rules(rules, num_rules, gg::kMaxRulesPerKind);
  g.L = num_labels;
  g.n = n;
  g.src.assign(src, src + m);
  g.dst.assign(dst, dst + m);
  g.lab.assign(lab, lab + m);
  b->graphs.push_back(std::move(g));
  if (graph_id) *graph_id = int(b->graphs.size() - 1);
  return GG_OK;
}

int gg_batch_run(gg_batch* b, uint64_t max_vertices, int keep, gg_stats* stats) {
  if (!b) return GG_E_ARG;
  std::string* cerr = nullptr;
  const int s = batch_run(b, max_vertices, keep, stats, cerr);
  // the context's last error names what failed (gg_last_error)
  if (s != GG_OK && cerr) *cerr = b->err;
  return s;
}

int gg_batch_info(const gg_batch* b, int* ng, int* ngroups) {
  if (!b) return GG_E_ARG;
  if (ng) *ng = int(b->graphs.size());
  if (ngroups) *ngroups = b->groups;
  return GG_OK;
}

int gg_batch_result(const gg_batch* b, int id, uint64_t* per_label, uint64_t* total, uint64_t* checksum) {
  if (!b || id < 0 || size_t(id) >= b->graphs.size()) return GG_E_ARG;
  const auto& g = b->graphs[size_t(id)];
  if (!g.done) return GG_E_STATE;
  if (per_label)
    for (int l = 0; l < g.L; ++l) per_label[l] = g.counts[size_t(l)];
  if (total) *total = g.total;
  if (checksum) *checksum = g.checksum;
  return GG_OK;
}

int gg_batch_export(const gg_batch* b, int id, uint32_t* buf, uint64_t cap, uint64_t* count) {
  if (!b || id < 0 || size_t(id) >= b->graphs.size()) return GG_E_ARG;
  const auto& g = b->graphs[size_t(id)];
  if (!g.done || g.triples.size() != size_t(g.total) * 3) return GG_E_STATE;
  if (count) *count = g.total;
  if (!buf || cap < g.total) return GG_E_ARG;
  std::copy(g.triples.begin(), g.triples.end(), buf);
  return GG_OK;
}

}  // extern "C"
