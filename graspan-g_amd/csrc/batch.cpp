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
#include <memory>
#include <mutex>
#include <new>
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
  // the two lanes' engines (GG_BATCH_LANES=2) and the device region they
  // carve their buffers from: made by the first run, kept for the next run,
  // freed with the batch
  std::vector<gg::Engine*> lanes;
  std::shared_ptr<gg::DeviceRegion> region;
  ~gg_batch() {
    for (auto* e : lanes) gg::engine_destroy(e);
    region.reset();
  }
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
  a.group_rows += b.group_rows;
  a.group_deferred += b.group_deferred;
}

// Launch groups are planned one at a time, each when a lane is free to
// build it, so that the vertex cap of a group uses the device bytes per
// vertex of the groups that have finished. With two lanes (each an engine
// with its own HIP stream) one lane's fixpoint runs while the other reads
// back its last group's results, builds its next group on the host and
// uploads it; the fixpoints themselves take turns (one device-wide launch
// sequence already fills the GPU).
struct Planner {
  gg_batch* b;
  uint64_t max_vertices;
  uint64_t mem;        // device bytes a group may plan against (0: use kFirstGroupVertices)
  double share;        // of mem, per group
  std::vector<size_t> ord;
  size_t next = 0;     // first graph (in ord) not yet planned
  size_t cap;          // graphs per group at most (halved after a group ran out of memory)
  std::vector<std::pair<size_t, size_t>> redo;  // halves of groups that ran out of memory
  std::map<std::pair<int, std::vector<int32_t>>, double> bytes_per_vertex;
  static constexpr uint64_t kFirstGroupVertices = 2u << 20;

  uint64_t vertex_cap(const std::pair<int, std::vector<int32_t>>& key) const {
    if (max_vertices) return max_vertices;
    auto it = bytes_per_vertex.find(key);
    if (it == bytes_per_vertex.end() || !mem) return kFirstGroupVertices;
    return std::max<uint64_t>(1, uint64_t(double(mem) * share / std::max(it->second, 1.0)));
  }

  // The next group [i, j) of ord; false when every graph is planned. A
  // group's combined grammar must stay within the engine's limits: at most
  // kMaxLabels labels, kMaxRulesPerKind unary and binary rules (eps rules
  // become loops), and labels * vertices and pair slots * vertices below
  // 2^32. A grammar that would pass one of them starts a new group.
  int take(size_t& i0, size_t& j0, std::string& err) {
    if (!redo.empty()) {
      i0 = redo.back().first;
      j0 = redo.back().second;
      redo.pop_back();
      return 1;
    }
    const size_t G = ord.size();
    if (next >= G) return 0;
    const size_t i = next;
    std::map<std::pair<int, std::vector<int32_t>>, uint32_t> seen;
    uint32_t labels = 0;
    int unary = 0, binary = 0, slots = 0;
    uint64_t verts = 0;
    uint64_t vcap = ~0ull;  // the smallest vertex cap of the group's grammars
    size_t j = i;
    for (; j < G && j - i < cap; ++j) {
      const auto& g = b->graphs[ord[j]];
      auto key = std::make_pair(g.L, g.rules);
      const bool fresh = seen.find(key) == seen.end();
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
      if (hard) {
        err = "graph " + std::to_string(ord[j]) + " alone exceeds the launch limits";
        return -1;
      }
      if (fresh) {
        vcap = ncap;
        seen[key] = labels;
        labels = nl;
        unary = nu;
        binary = nb;
        slots = ns;
      }
      verts = nv;
    }
    next = j;
    i0 = i;
    j0 = j;
    return 1;
  }
};

// Builds, uploads, closes and reads back the launch group ord[i..j) on one
// engine. GG_E_NOMEM is returned to the caller (which splits the group).
int run_group(gg_batch* b, gg::Engine* e, const std::vector<size_t>& ord, size_t i, size_t j, int keep,
              gg_stats& one, double& bytes_per_vertex, bool trace, int lane, std::mutex* turn, std::string& err) {
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto ms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point c) {
    return std::chrono::duration<double, std::milli>(c - a).count();
  };
  const auto t_start = now();
  // combined grammar: each distinct grammar (in order of first appearance in
  // the group) gets the next range of label ids
  std::map<std::pair<int, std::vector<int32_t>>, uint32_t> grammar_base;
  std::vector<std::pair<int, std::vector<int32_t>>> order;
  uint32_t labels = 0;
  uint64_t verts = 0;
  for (size_t k = i; k < j; ++k) {
    const auto& g = b->graphs[ord[k]];
    auto key = std::make_pair(g.L, g.rules);
    if (grammar_base.emplace(key, labels).second) {
      labels += uint32_t(g.L);
      order.push_back(key);
    }
    verts += g.n;
  }
  std::vector<int32_t> rules;
  for (auto& key : order) {
    const uint32_t lb = grammar_base[key];
    const auto& r = key.second;
    for (size_t k = 0; k + 2 < r.size(); k += 3) {
      if (r[k + 1] == 0) continue;  // eps: explicit loops below
      rules.push_back(r[k] + int32_t(lb));
      rules.push_back(r[k + 1] + int32_t(lb));
      rules.push_back(r[k + 2] ? r[k + 2] + int32_t(lb) : 0);
    }
  }
  gg::Grammar cg;
  int s = gg::grammar_from_rules(int(rules.size() / 3), rules.data(), int(std::max(labels, 1u)), cg, err);
  if (s != GG_OK) return s;
  // super graph: every graph's edges shifted to its vertex range, plus the
  // eps loops its grammar asks for (sized first, then filled in place)
  gg::Graph sg;
  sg.n = uint32_t(verts);
  // each graph's slice of the union: its vertex base, label base and first
  // position, then the slices are filled by a few host threads
  std::vector<uint32_t> v0(1, 0), l0, lg;
  std::vector<uint64_t> at0;
  uint64_t at = 0;
  for (size_t k = i; k < j; ++k) {
    const auto& g = b->graphs[ord[k]];
    at0.push_back(at);
    at += g.src.size();
    for (size_t r = 0; r + 2 < g.rules.size(); r += 3)
      if (g.rules[r + 1] == 0) at += g.n;
    l0.push_back(grammar_base[std::make_pair(g.L, g.rules)]);
    lg.push_back(uint32_t(g.L));
    v0.push_back(v0.back() + g.n);
  }
  sg.src.resize(at);
  sg.dst.resize(at);
  sg.lab.resize(at);
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
  const size_t nt = std::min<size_t>({ng, 8, std::max(1u, std::thread::hardware_concurrency())});
  if (nt <= 1 || at < (1u << 20)) {
    fill(0, ng);
  } else {
    std::vector<std::thread> pool;
    for (size_t t = 0; t < nt; ++t) pool.emplace_back(fill, ng * t / nt, ng * (t + 1) / nt);
    for (auto& th : pool) th.join();
  }
  const auto t_built = now();
  if ((s = gg::engine_set_grammar(e, cg, err)) != GG_OK) return s;
  const uint64_t group_m = sg.src.size();
  s = gg::engine_set_graph(e, std::move(sg), err);
  const auto t_up = now();
  one = gg_stats{};
  // lanes take turns at the fixpoint; reading results, building and
  // uploading overlap the other lane's turn. A fixpoint hands the turn on
  // once its rounds have few active rows (its tail, which leaves most of the
  // device idle): the next lane's first rounds overlap it. Measured on the
  // 4096-graph set: 1/8 of the rows 9.75 s, 1/16 9.94 s, never 9.98 s, 1/4
  // 10.5 s (two closures at their peak no longer fit, groups split).
  std::unique_lock<std::mutex> lk;
  if (turn && s == GG_OK) lk = std::unique_lock<std::mutex>(*turn);
  const auto t_turn = now();
  if (lk.owns_lock()) {
    // GG_BATCH_HANDOFF=d: hand on below rows / d active rows (default 8)
    uint64_t d = 8;
    if (const char* h = std::getenv("GG_BATCH_HANDOFF")) d = std::max<uint64_t>(1, std::strtoull(h, nullptr, 10));
    gg::engine_on_round(e, [&lk, d](uint64_t active, uint64_t rows) {
      if (lk.owns_lock() && active * d < rows) lk.unlock();
    });
  }
  if (s == GG_OK) s = gg::engine_run(e, 0, &one, err);
  gg::engine_on_round(e, nullptr);
  if (lk.owns_lock()) lk.unlock();
  const auto t_run = now();
  if (s != GG_OK) {
    if (trace && s == GG_E_NOMEM)
      std::fprintf(stderr, "[gg-batch] lane %d: group of %zu graphs, %llu vertices: out of device memory after %.0f ms\n",
                   lane, j - i, (unsigned long long)verts, ms(t_start, t_run));
    gg::engine_release(e);
    return s;
  }
  bytes_per_vertex = double(one.device_bytes_peak) / double(verts);
  std::vector<uint64_t> counts, sums;
  if ((s = gg::engine_slices(e, v0, l0, lg, counts, sums, err)) != GG_OK) return s;
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
                                       &cnt, err)) != GG_OK)
        return s;
    }
    g.done = true;
  }
  if (trace)
    std::fprintf(stderr,
                 "[gg-batch] lane %d: %zu graphs, %llu vertices, %u labels, %llu edges -> %llu closure edges, "
                 "%u rounds; build %.0f ms, upload %.0f ms, wait %.0f ms, run %.0f ms, results %.0f ms, "
                 "device peak %.1f GB\n",
                 lane, j - i, (unsigned long long)verts, labels, (unsigned long long)group_m,
                 (unsigned long long)one.total_edges, one.rounds, ms(t_start, t_built), ms(t_built, t_up),
                 ms(t_up, t_turn), ms(t_turn, t_run), ms(t_run, now()), double(one.device_bytes_peak) / 1e9);
  return GG_OK;
}

int batch_run(gg_batch* b, uint64_t max_vertices, int keep, gg_stats* stats, std::string*& cerr) {
  int st = GG_OK;
  gg::Engine* e = gg::ctx_engine(b->ctx, st, cerr);
  if (!e) return fail(b, st, cerr ? *cerr : "no device");
  b->groups = 0;
  const size_t G = b->graphs.size();
  Planner pl{b, max_vertices, max_vertices ? 0 : gg::ctx_memory(b->ctx), gg::tuned_group_mem_share(), {}, 0, G,
             {}, {}};
  // Graphs are ordered by grammar before they are cut into launch groups: a
  // group's rows are its vertices times all its labels, so a group that
  // mixes fewer grammars carries fewer empty rows through every round.
  pl.ord.resize(G);
  for (size_t k = 0; k < G; ++k) pl.ord[k] = k;
  std::stable_sort(pl.ord.begin(), pl.ord.end(), [&](size_t x, size_t y) {
    const auto& gx = b->graphs[x];
    const auto& gy = b->graphs[y];
    return std::make_pair(gx.L, gx.rules) < std::make_pair(gy.L, gy.rules);
  });
  // GG_BATCH_TRACE=1: one stderr line per launch group (sizes and the host
  // time of each phase)
  const char* tr = std::getenv("GG_BATCH_TRACE");
  const bool trace = tr && tr[0] == '1';
  // GG_BATCH_LANES: launch groups in flight at once (default 2; 1 runs the
  // groups one after another on the context's own engine, host work and
  // device work in turn)
  int lanes = 2;
  if (const char* ln = std::getenv("GG_BATCH_LANES")) lanes = std::max(1, std::min(2, std::atoi(ln)));
  if (G < 2) lanes = 1;
  // Two lanes carve their buffers from one device region, allocated once
  // (per-buffer hipMalloc/hipFree of tens of GB would synchronise the lanes
  // and costs seconds once memory has been freed before). A lane frees its
  // closure as soon as its results are read, so the lane whose fixpoint runs
  // may use nearly all of the region: groups are planned as for one lane.
  std::vector<gg::Engine*> eng{e};
  if (lanes > 1 && b->lanes.empty()) {
    std::string err;
    const int dev = gg::engine_device(e);
    // the context's engine gives its cached blocks back to the device
    gg::engine_release(e);
    gg::engine_trim(e);
    uint64_t total = 0, free_bytes = 0;
    if (gg::device_query(nullptr, dev, nullptr, nullptr, 0, &total, &free_bytes, err) == GG_OK) {
      uint64_t bytes = uint64_t(double(free_bytes) * 0.92) & ~uint64_t((2u << 20) - 1);
      const uint64_t mem = gg::ctx_memory(b->ctx);
      if (mem) bytes = std::min<uint64_t>(bytes, mem);
      b->region = gg::device_region_create(dev, bytes, err);
    }
    for (int l = 0; l < lanes && b->region; ++l) {
      int s = GG_OK;
      gg::Engine* x = gg::engine_create(dev, 0, 0, s, err);
      if (!x) return fail(b, s, err);
      b->lanes.push_back(x);
      if ((s = gg::engine_use_region(x, b->region, err)) != GG_OK) return fail(b, s, err);
    }
  }
  // per-graph results are read from the packed entries: no id expansion of
  // the packed labels after each group's last round (exports write them)
  for (auto* x : eng) gg::engine_defer_unpack(x, true);
  for (auto* x : b->lanes) gg::engine_defer_unpack(x, true);
  if (lanes > 1 && b->lanes.size() == size_t(lanes)) {
    eng = b->lanes;
    // the same group sizes as one lane on the whole device
    if (pl.mem) {
      const uint64_t bytes = gg::region_bytes(b->region.get());
      pl.share = std::min(0.9, pl.share * double(pl.mem) / double(bytes));
      pl.mem = bytes;
    }
  } else {
    lanes = 1;
  }
  std::mutex mu, turn;
  gg_stats total{};
  int first_err = GG_OK;
  std::string err_msg;
  std::vector<std::pair<size_t, size_t>> alone;  // lanes: single graphs to run again alone
  // GG_BATCH_TURNS=0: the lanes' fixpoints run at the same time (no turns)
  bool take_turns = true;
  if (const char* tt = std::getenv("GG_BATCH_TURNS")) take_turns = tt[0] != '0';
  auto worker = [&](int lane) {
    std::string err;
    for (;;) {
      size_t i = 0, j = 0;
      {
        std::lock_guard<std::mutex> lk(mu);
        if (first_err != GG_OK) return;
        const int t = pl.take(i, j, err);
        if (t < 0) {
          first_err = GG_E_UNSUPPORTED;
          err_msg = err;
          return;
        }
        if (t == 0) return;
      }
      gg_stats one{};
      double bpv = 0;
      int s = GG_OK;
      try {
        s = run_group(b, eng[size_t(lane)], pl.ord, i, j, keep, one, bpv, trace, lane,
                      lanes > 1 && take_turns ? &turn : nullptr, err);
      } catch (const std::bad_alloc&) {
        // host memory for the union graph or the exported triples: nothing
        // may escape a worker thread (or the C ABI)
        s = GG_E_NOMEM;
        err = "host memory exhausted building or reading a launch group";
      }
      if (lanes > 1) gg::engine_release(eng[size_t(lane)]);  // the closure's memory, for the other lane
      std::lock_guard<std::mutex> lk(mu);
      if (s == GG_E_NOMEM && j - i > 1) {
        // the closure size was not known in advance: cut the group in half
        // and run both halves again. One lane: later groups keep the smaller
        // count. Two lanes: the other lane's closure may have been on the
        // device too, so the count stays (the bytes per vertex of the groups
        // that finish size the next ones).
        const size_t mid = i + (j - i) / 2;
        if (lanes == 1) pl.cap = std::min(pl.cap, mid - i);
        pl.redo.push_back({mid, j});
        pl.redo.push_back({i, mid});
        continue;
      }
      if (s == GG_E_NOMEM && lanes > 1) {
        // one graph that did not fit beside the other lane's closure runs
        // again alone once both lanes are done
        alone.push_back({i, j});
        continue;
      }
      if (s != GG_OK) {
        if (first_err == GG_OK) {
          first_err = s;
          err_msg = err;
        }
        return;
      }
      add_stats(total, one);
      // a group of one grammar sets that grammar's bytes per vertex; a mixed
      // group's figure is an average over its grammars, which may be far
      // below the densest one's: it only raises figures already known
      std::map<std::pair<int, std::vector<int32_t>>, int> kinds;
      for (size_t k = i; k < j; ++k) {
        const auto& g = b->graphs[pl.ord[k]];
        kinds[std::make_pair(g.L, g.rules)] = 1;
      }
      for (auto& kv : kinds) {
        auto it = pl.bytes_per_vertex.find(kv.first);
        if (kinds.size() == 1)
          pl.bytes_per_vertex[kv.first] = bpv;
        else if (it != pl.bytes_per_vertex.end())
          it->second = std::max(it->second, bpv);
      }
      ++b->groups;
    }
  };
  if (lanes == 1) {
    worker(0);
  } else {
    std::vector<std::thread> th;
    for (int l = 0; l < lanes; ++l) th.emplace_back(worker, l);
    for (auto& t : th) t.join();
    // single graphs that ran out of memory beside the other lane, now with
    // the whole region
    for (size_t k = 0; k < alone.size() && first_err == GG_OK; ++k) {
      std::string err;
      gg_stats one{};
      double bpv = 0;
      const int s = run_group(b, eng[0], pl.ord, alone[k].first, alone[k].second, keep, one, bpv, trace, 0, nullptr,
                              err);
      gg::engine_release(eng[0]);
      if (s != GG_OK) {
        first_err = s;
        err_msg = err;
        break;
      }
      add_stats(total, one);
      ++b->groups;
    }
  }
  gg::engine_defer_unpack(e, false);  // the context's own runs write every id
  if (first_err != GG_OK) return fail(b, first_err, err_msg);
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
