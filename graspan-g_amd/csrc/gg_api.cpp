// C ABI of the closure engine (declarations and reference citations in
// include/graspan_g.h). Host C++; the device work lives in closure.hip.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <utility>
#include <vector>
#include <string>

#include "engine.h"
#include "formats.h"
#include "graspan_g.h"
#include "hash.h"

struct gg_ctx {
  gg_options opts{-1, 0, 0};
  std::string err;
  gg::Grammar grammar;
  bool has_grammar = false;
  gg::Graph graph;
  bool has_graph = false;
  bool graph_on_device = false;
  gg::Engine* engine = nullptr;
};

namespace {

int fail(gg_ctx* c, int status, const std::string& msg) {
  c->err = msg;
  return status;
}

// Row ids (label * vertices + vertex) are 32-bit in every schedule: a graph
// with num_labels * num_vertices >= 2^32 is refused before any device work.
int check_row_ids(gg_ctx* c) {
  if (uint64_t(c->grammar.num_labels()) * c->graph.n >= 0xFFFFFFFFull)
    return fail(c, GG_E_UNSUPPORTED, "num_labels * num_vertices must be < 2^32 (row ids are 32-bit)");
  return GG_OK;
}

// Creates the engine on first use (the only place that touches the GPU).
int ensure_engine(gg_ctx* c) {
  if (c->engine) return GG_OK;
  int status = GG_OK;
  c->engine = gg::engine_create(c->opts.device, c->opts.device_budget, c->opts.verbose, status, c->err);
  if (!c->engine) return status;
  if (c->has_grammar) {
    int s = gg::engine_set_grammar(c->engine, c->grammar, c->err);
    if (s != GG_OK) return s;
  }
  return GG_OK;
}

int push_graph(gg_ctx* c) {
  int s = ensure_engine(c);
  if (s != GG_OK) return s;
  s = gg::engine_set_graph(c->engine, c->graph, c->err);
  if (s == GG_OK) c->graph_on_device = true;
  return s;
}

// The reference's vertex intervals: initializePartitions (src/main.cpp:141-175)
// over the degree pre-pass of getGraphInfo (src/main.cpp:106-139), which
// counts per source its input lines and its distinct (dst/192, label)
// "elements" (192-id bitmap elements, src/kernel.h:35-41). The heap is the
// reference's fixed DEVMEMUSED_MB = 10000 (src/kernel.h:26, src/kernel.cu:35);
// a partition may hold heap/4 words of 8*elements + roundup32(vertices)*labels.
// One partition that fits is split at the element midpoint.
// GG_REF_HEAP_WORDS replaces the heap size (in 32-bit words), so that tests
// can reach the multi-partition cut loop (src/main.cpp:145-155) with small
// graphs.
uint64_t reference_heap_words() {
  if (const char* h = std::getenv("GG_REF_HEAP_WORDS")) {
    const uint64_t w = std::strtoull(h, nullptr, 10);
    if (w) return w;
  }
  return (10000ull / 8 * 7) * 1024 * 256;  // allocateElementPool, src/kernel.cu:35
}

int reference_partitions(const gg::Graph& gr, int L, std::vector<std::pair<uint32_t, uint32_t>>& parts,
                         std::string& err) {
  const uint32_t n = gr.n;
  std::vector<uint64_t> elems(n, 0);
  {
    std::vector<std::vector<uint64_t>> keys(n);
    for (size_t i = 0; i < gr.src.size(); ++i) keys[gr.src[i]].push_back((uint64_t(gr.dst[i] / 192) << 8) | gr.lab[i]);
    for (uint32_t u = 0; u < n; ++u) {
      auto& k = keys[u];
      std::sort(k.begin(), k.end());
      elems[u] = uint64_t(std::unique(k.begin(), k.end()) - k.begin());
    }
  }
  const uint64_t limit = reference_heap_words() / 4;
  auto rup32 = [](uint64_t x) { return (x + 31) / 32 * 32; };
  parts.clear();
  uint64_t t1 = 0, t3 = 0;
  for (uint32_t i = 0; i < n; ++i) {
    t1 += elems[i];
    if (t1 * 8 + rup32(i - t3 + 1) * uint64_t(L) > limit) {
      if (i == 0 || uint64_t(i - 1) < t3) {
        err = "degenerate partitioning";
        return GG_E_UNSUPPORTED;
      }
      parts.push_back({uint32_t(t3), i - 1});
      t1 = elems[i];
      t3 = i;
    }
  }
  parts.push_back({uint32_t(t3), n - 1});
  if (parts.size() == 1) {
    uint64_t t4 = 0;
    for (uint32_t i = 0; i < n; ++i) {
      t4 += elems[i];
      if (t4 > t1 / 2) {
        if (i == 0) break;
        parts[0].second = i - 1;
        parts.push_back({i, n - 1});
        return GG_OK;
      }
    }
    err = "degenerate partitioning: the reference's small-graph split needs a second vertex interval";
    return GG_E_UNSUPPORTED;
  }
  return GG_OK;
}

}  // namespace

namespace gg {
RuleKinds rule_kinds(const int32_t* rules, int num_rules) {
  RuleKinds k;
  std::vector<int32_t> derived;  // lhs of unary and binary rules
  for (int i = 0; i < num_rules; ++i) {
    const int32_t* r = rules + 3 * i;
    if (r[1] == 0) {
      ++k.eps;
    } else {
      derived.push_back(r[0]);
      if (r[2] == 0)
        ++k.unary;
      else
        ++k.binary;
    }
  }
  for (int i = 0; i < num_rules; ++i) {
    const int32_t* r = rules + 3 * i;
    if (r[1] != 0 && r[2] != 0 && std::find(derived.begin(), derived.end(), r[2]) != derived.end()) ++k.slots;
  }
  return k;
}

// The engine of a context, created on first use (batch.cpp).
Engine* ctx_engine(gg_ctx* c, int& status, std::string*& err) {
  err = &c->err;
  status = ensure_engine(c);
  if (status != GG_OK) return nullptr;
  c->graph_on_device = false;  // a batch replaces the context's device graph
  return c->engine;
}

// Device memory a batch may plan for: the context's budget, else the
// device's total memory (0 if it cannot be read).
uint64_t ctx_memory(gg_ctx* c) {
  if (c->opts.device_budget) return c->opts.device_budget;
  uint64_t total = 0, free_bytes = 0;
  std::string err;
  if (device_query(nullptr, c->opts.device, nullptr, nullptr, 0, &total, &free_bytes, err) != GG_OK) return 0;
  return total;
}
}  // namespace gg

extern "C" {

const char* gg_version(void) { return "graspan-g-amd 0.1 (gfx950)"; }

int gg_ctx_create(const gg_options* opts, gg_ctx** out) {
  if (!out) return GG_E_ARG;
  gg_ctx* c = new (std::nothrow) gg_ctx();
  if (!c) return GG_E_NOMEM;
  if (opts) c->opts = *opts;
  *out = c;
  return GG_OK;
}

int gg_ctx_destroy(gg_ctx* c) {
  if (!c) return GG_E_ARG;
  if (c->engine) gg::engine_destroy(c->engine);
  delete c;
  return GG_OK;
}

const char* gg_last_error(const gg_ctx* c) { return c ? c->err.c_str() : "null context"; }

int gg_grammar_load_file(gg_ctx* c, const char* path) {
  if (!c || !path) return c ? fail(c, GG_E_ARG, "null path") : GG_E_ARG;
  gg::Grammar g;
  int s = gg::parse_grammar_file(path, g, c->err);
  if (s != GG_OK) return s;
  c->grammar = std::move(g);
  c->has_grammar = true;
  c->has_graph = false;  // label ids changed: the graph must be reloaded
  c->graph_on_device = false;
  if (c->engine) return gg::engine_set_grammar(c->engine, c->grammar, c->err);
  return GG_OK;
}

int gg_grammar_set(gg_ctx* c, int num_rules, const int32_t* rules, int num_labels) {
  if (!c) return GG_E_ARG;
  gg::Grammar g;
  int s = gg::grammar_from_rules(num_rules, rules, num_labels, g, c->err);
  if (s != GG_OK) return s;
  c->grammar = std::move(g);
  c->has_grammar = true;
  c->has_graph = false;
  c->graph_on_device = false;
  if (c->engine) return gg::engine_set_grammar(c->engine, c->grammar, c->err);
  return GG_OK;
}

int gg_grammar_info(const gg_ctx* c, int* num_rules, int* num_labels) {
  if (!c) return GG_E_ARG;
  if (!c->has_grammar) return GG_E_STATE;
  if (num_rules) *num_rules = int(c->grammar.rules.size());
  if (num_labels) *num_labels = c->grammar.num_labels();
  return GG_OK;
}

int gg_grammar_rules(const gg_ctx* c, int32_t* out) {
  if (!c || !out) return GG_E_ARG;
  if (!c->has_grammar) return GG_E_STATE;
  for (size_t i = 0; i < c->grammar.rules.size(); ++i)
    for (int j = 0; j < 3; ++j) out[3 * i + j] = c->grammar.rules[i][j];
  return GG_OK;
}

int gg_grammar_plan_fixpoint(int num_rules, const int32_t* rules, int num_labels, int32_t* out, int cap,
                             int* out_rules) {
  gg::Grammar tmp;
  std::string err;
  if (!out_rules || (cap > 0 && !out) || gg::grammar_from_rules(num_rules, rules, num_labels, tmp, err) != GG_OK)
    return GG_E_ARG;
  const std::vector<int32_t> p = gg::plan_fixpoint_rules(rules, num_rules, gg::kMaxRulesPerKind);
  *out_rules = int(p.size() / 3);
  if (*out_rules > cap) return GG_E_ARG;
  std::copy(p.begin(), p.end(), out);
  return GG_OK;
}

const char* gg_label_name(const gg_ctx* c, int id) {
  if (!c || !c->has_grammar || id < 1 || id > c->grammar.num_labels()) return nullptr;
  return c->grammar.names[size_t(id - 1)].c_str();
}

int gg_label_id(const gg_ctx* c, const char* name) {
  if (!c || !name || !c->has_grammar) return 0;
  auto it = c->grammar.ids.find(name);
  return it == c->grammar.ids.end() ? 0 : it->second;
}

int gg_graph_load_file(gg_ctx* c, const char* path) {
  if (!c || !path) return c ? fail(c, GG_E_ARG, "null path") : GG_E_ARG;
  if (!c->has_grammar) return fail(c, GG_E_STATE, "load a grammar before the graph (labels are resolved by name)");
  gg::Graph g;
  int s = gg::parse_graph_file(path, c->grammar, g, c->err);
  if (s != GG_OK) return s;
  if (g.n == 0) return fail(c, GG_E_FORMAT, "graph has no vertices");
  c->graph = std::move(g);
  c->has_graph = true;
  c->graph_on_device = false;
  if (c->engine) return push_graph(c);
  return GG_OK;
}

int gg_graph_set_coo(gg_ctx* c, uint32_t nv, uint64_t m, const uint32_t* src, const uint32_t* dst, const uint8_t* lab) {
  if (!c) return GG_E_ARG;
  if (!c->has_grammar) return fail(c, GG_E_STATE, "set a grammar before the graph");
  if (nv == 0) return fail(c, GG_E_ARG, "num_vertices must be > 0");
  if (m && (!src || !dst || !lab)) return fail(c, GG_E_ARG, "null edge array");
  const int L = c->grammar.num_labels();
  for (uint64_t i = 0; i < m; ++i) {
    if (src[i] >= nv || dst[i] >= nv)
      return fail(c, GG_E_ARG, "edge " + std::to_string(i) + ": vertex id out of range");
    if (lab[i] < 1 || lab[i] > L) return fail(c, GG_E_ARG, "edge " + std::to_string(i) + ": label id out of range");
  }
  gg::Graph g;
  g.n = nv;
  g.src.assign(src, src + m);
  g.dst.assign(dst, dst + m);
  g.lab.assign(lab, lab + m);
  c->graph = std::move(g);
  c->has_graph = true;
  c->graph_on_device = false;
  if (c->engine) return push_graph(c);
  return GG_OK;
}

int gg_graph_info(const gg_ctx* c, uint32_t* nv, uint64_t* m) {
  if (!c) return GG_E_ARG;
  if (!c->has_graph) return GG_E_STATE;
  if (nv) *nv = c->graph.n;
  if (m) *m = c->graph.m();
  return GG_OK;
}

int gg_graph_edges(const gg_ctx* c, uint32_t* src, uint32_t* dst, uint8_t* lab) {
  if (!c) return GG_E_ARG;
  if (!c->has_graph) return GG_E_STATE;
  const size_t m = c->graph.m();
  if (src) std::memcpy(src, c->graph.src.data(), m * 4);
  if (dst) std::memcpy(dst, c->graph.dst.data(), m * 4);
  if (lab) std::memcpy(lab, c->graph.lab.data(), m);
  return GG_OK;
}

int gg_closure_run_partitioned(gg_ctx* c, int filter_flag, gg_stats* stats) {
  if (!c) return GG_E_ARG;
  if (!c->has_grammar || !c->has_graph) return fail(c, GG_E_STATE, "load a grammar and a graph first");
  if (check_row_ids(c) != GG_OK) return GG_E_UNSUPPORTED;
  std::vector<std::pair<uint32_t, uint32_t>> parts;
  int s = reference_partitions(c->graph, c->grammar.num_labels(), parts, c->err);
  if (s != GG_OK) return s;
  s = ensure_engine(c);
  if (s != GG_OK) return s;
  if (!c->graph_on_device) {
    s = push_graph(c);
    if (s != GG_OK) return s;
  }
  return gg::engine_run_partitioned(c->engine, filter_flag, parts, gg::PartOpts{}, stats, c->err);
}

int gg_closure_run_out_of_core(gg_ctx* c, int num_partitions, gg_stats* stats) {
  if (!c) return GG_E_ARG;
  if (!c->has_grammar || !c->has_graph) return fail(c, GG_E_STATE, "load a grammar and a graph first");
  if (check_row_ids(c) != GG_OK) return GG_E_UNSUPPORTED;
  const uint32_t n = c->graph.n;
  if (num_partitions < 1 || uint32_t(num_partitions) > n) return fail(c, GG_E_ARG, "bad partition count");
  std::vector<std::pair<uint32_t, uint32_t>> parts;
  for (int p = 0; p < num_partitions; ++p) {
    const uint32_t a = uint32_t(uint64_t(n) * uint64_t(p) / uint64_t(num_partitions));
    const uint32_t b = uint32_t(uint64_t(n) * uint64_t(p + 1) / uint64_t(num_partitions));
    if (b > a) parts.push_back({a, b - 1});
  }
  int s = ensure_engine(c);
  if (s != GG_OK) return s;
  if (!c->graph_on_device) {
    s = push_graph(c);
    if (s != GG_OK) return s;
  }
  gg::PartOpts po;
  po.spill = true;
  po.ddm = true;
  return gg::engine_run_partitioned(c->engine, 0, parts, po, stats, c->err);
}

int gg_reference_partitions(const gg_ctx* c, uint32_t* first, uint32_t* last, int cap, int* count) {
  if (!c || !count) return GG_E_ARG;
  if (!c->has_grammar || !c->has_graph) return GG_E_STATE;
  std::vector<std::pair<uint32_t, uint32_t>> parts;
  std::string err;
  const int s = reference_partitions(c->graph, c->grammar.num_labels(), parts, err);
  if (s != GG_OK) return s;
  *count = int(parts.size());
  for (int i = 0; i < cap && size_t(i) < parts.size(); ++i) {
    if (first) first[i] = parts[size_t(i)].first;
    if (last) last[i] = parts[size_t(i)].second;
  }
  return GG_OK;
}

int gg_closure_run(gg_ctx* c, int filter_flag, gg_stats* stats) {
  if (!c) return GG_E_ARG;
  if (!c->has_grammar || !c->has_graph) return fail(c, GG_E_STATE, "load a grammar and a graph first");
  if (filter_flag != 0) return gg_closure_run_partitioned(c, filter_flag, stats);
  if (check_row_ids(c) != GG_OK) return GG_E_UNSUPPORTED;
  int s = ensure_engine(c);
  if (s != GG_OK) return s;
  if (!c->graph_on_device) {
    s = push_graph(c);
    if (s != GG_OK) return s;
  }
  return gg::engine_run(c->engine, filter_flag, stats, c->err);
}

int gg_closure_counts(gg_ctx* c, uint64_t* per_label, uint64_t* total) {
  if (!c) return GG_E_ARG;
  if (!c->engine) return fail(c, GG_E_STATE, "no closure: call gg_closure_run first");
  return gg::engine_counts(c->engine, per_label, total, c->err);
}

int gg_closure_export(gg_ctx* c, uint32_t* buf, uint64_t cap, uint64_t* count) {
  if (!c) return GG_E_ARG;
  if (!c->engine) return fail(c, GG_E_STATE, "no closure: call gg_closure_run first");
  return gg::engine_export(c->engine, buf, cap, count, c->err);
}

int gg_closure_degrees(gg_ctx* c, uint64_t* deg) {
  if (!c || !deg) return GG_E_ARG;
  if (!c->engine) return fail(c, GG_E_STATE, "no closure: call gg_closure_run first");
  return gg::engine_degrees(c->engine, deg, c->err);
}

int gg_closure_checksum(gg_ctx* c, uint64_t* sum) {
  if (!c || !sum) return GG_E_ARG;
  if (!c->engine) return fail(c, GG_E_STATE, "no closure: call gg_closure_run first");
  return gg::engine_checksum(c->engine, sum, c->err);
}

uint64_t gg_triple_hash(uint32_t s, uint32_t d, uint32_t l) { return gg::triple_hash(s, d, l); }

int gg_closure_release(gg_ctx* c) {
  if (!c) return GG_E_ARG;
  if (c->engine) gg::engine_release(c->engine);
  return GG_OK;
}

int gg_closure_run_noalloc(int num_rules, const int32_t* rules, int num_labels, uint32_t num_vertices,
                           uint64_t num_edges, const uint32_t* src, const uint32_t* dst, const uint8_t* label,
                           void* stream, uint64_t* row_ptr, uint32_t* col, uint64_t col_cap, uint64_t* total,
                           uint64_t* checksum, uint64_t* per_label, char* err, size_t err_cap) {
  std::string msg;
  int s = GG_OK;
  gg::Grammar g;
  if (!row_ptr || (col_cap && !col) || num_vertices == 0 || (num_edges && (!src || !dst || !label))) {
    s = GG_E_ARG;
    msg = "null buffer or no vertices";
  } else {
    s = gg::grammar_from_rules(num_rules, rules, num_labels, g, msg);
  }
  if (s == GG_OK)
    s = gg::engine_run_noalloc(g, num_vertices, num_edges, src, dst, label, stream, row_ptr, col, col_cap, total,
                               checksum, per_label, msg);
  if (err && err_cap) {
    const size_t k = std::min(err_cap - 1, s == GG_OK ? size_t(0) : msg.size());
    std::memcpy(err, msg.data(), k);
    err[k] = 0;
  }
  return s;
}

int gg_device_info(gg_ctx* c, int* count, int* selected, char* name, size_t name_cap, uint64_t* total,
                   uint64_t* free_bytes) {
  if (!c) return GG_E_ARG;
  return gg::device_query(count, c->opts.device, selected, name, name_cap, total, free_bytes, c->err);
}

}  // extern "C"
