// Interface between the C ABI (gg_api.cpp) and the device engine
// (closure.hip).
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "formats.h"
#include "graspan_g.h"

namespace gg {

class Engine;

// The engine takes at most this many eps, unary and binary rules each.
constexpr int kMaxRulesPerKind = 64;

// Rules of each kind in a rule table (lhs, rhs1, rhs2 triples, 0 = absent)
// and its pair-list slots: binary rules whose second label is the lhs of a
// unary or binary rule. The engine holds one row per slot and vertex for the
// pair lists, so slots * vertices must stay below 2^32 like labels * vertices.
struct RuleKinds {
  int eps = 0, unary = 0, binary = 0, slots = 0;
};
RuleKinds rule_kinds(const int32_t* rules, int num_rules);

// Creates the device engine on `device` (-1 = most CUs). Returns nullptr and
// sets status/err on failure.
Engine* engine_create(int device, uint64_t budget, int verbose, int& status, std::string& err);
void engine_destroy(Engine* e);
// Hands the engine's cached device blocks back to the device.
void engine_trim(Engine* e);

// Uploads the grammar rule table and the input edges (host copies). The
// edges stay resident on the device across runs.
int engine_set_grammar(Engine* e, const Grammar& g, std::string& err);
int engine_set_graph(Engine* e, const Graph& g, std::string& err);
// Same, taking over the caller's host arrays instead of copying them.
int engine_set_graph(Engine* e, Graph&& g, std::string& err);

int engine_run(Engine* e, int filter, gg_stats* stats, std::string& err);
// Options of the partitioned schedule: spill = partitions not in the current
// session live in pinned host memory; ddm = run a partition pair only when
// edges cross between the two and one of them has new edges.
struct PartOpts {
  bool spill = false;
  bool ddm = false;
};
// The reference's partitioned schedule over the vertex intervals `parts`.
int engine_run_partitioned(Engine* e, int filter, const std::vector<std::pair<uint32_t, uint32_t>>& parts,
                           const PartOpts& po, gg_stats* stats, std::string& err);
int engine_counts(Engine* e, uint64_t* per_label, uint64_t* total, std::string& err);
int engine_export(Engine* e, uint32_t* buf, uint64_t cap, uint64_t* count, std::string& err);
int engine_degrees(Engine* e, uint64_t* deg, std::string& err);
int engine_checksum(Engine* e, uint64_t* sum, std::string& err);
// Per-graph results of a batched closure (graph g: vertices [v0[g], v0[g+1]),
// combined labels [l0[g], l0[g]+lg[g])): per-label counts (concatenated over
// graphs) and per-graph checksums in the graph's own numbering.
int engine_slices(Engine* e, const std::vector<uint32_t>& v0, const std::vector<uint32_t>& l0,
                  const std::vector<uint32_t>& lg, std::vector<uint64_t>& counts, std::vector<uint64_t>& sums,
                  std::string& err);
int engine_export_slice(Engine* e, uint32_t v0, uint32_t v1, uint32_t l0, uint32_t lg, uint32_t* buf, uint64_t cap,
                        uint64_t* count, std::string& err);
int engine_release(Engine* e);

// Sessions: one running fixpoint that graphs join at round boundaries and
// leave when their rows stop changing (the serving loop). A graph occupies
// vertices [v0, v0 + n) and the combined labels [l0, l0 + lg) (0-based) of
// the session's grammar, which must have no eps rules (they are explicit
// loop edges). engine_session_admit grows the vertex range to new_n and
// queues the given edges (session numbering) for the next round;
// engine_session_round runs one round and returns, per live graph, its edges
// in the next delta (0: it has reached its fixpoint); engine_session_results
// gives finished graphs' per-label counts (concatenated) and checksums in the
// graphs' own numbering. engine_session_end drops the session's state (trim:
// also hand the engine's cached device memory back).
struct SessionSlice {
  uint32_t v0, n, l0, lg;
};
int engine_session_begin(Engine* e, const Grammar& combined, std::string& err);
int engine_session_admit(Engine* e, uint32_t new_n, const uint32_t* src, const uint32_t* dst, const uint8_t* lab,
                         uint64_t m, std::string& err);
int engine_session_round(Engine* e, const std::vector<SessionSlice>& live, std::vector<uint64_t>& delta,
                         std::string& err);
int engine_session_results(Engine* e, const std::vector<SessionSlice>& done, std::vector<uint64_t>& counts,
                           std::vector<uint64_t>& sums, std::string& err);
void engine_session_end(Engine* e, bool trim);
// Device bytes the engine holds (in use plus its block cache).
uint64_t engine_device_bytes(Engine* e);
// Session statistics: rounds, P/K/O/N, kernel times by phase when timing is
// on (events around each phase of every round), device peak; alloc_ms gets
// the engine's host time spent allocating device memory since it was made.
void engine_session_timing(Engine* e, bool on);
// A device region allocated once and carved by best fit (freed when the last
// engine using it is gone). Engines given one carve every buffer from it
// instead of allocating per buffer, so long-running callers (the serving
// loop) avoid the per-round hipMalloc/hipFree of growing buffers; several
// engines may share one region. engine_use_region must come before the
// engine allocates anything.
struct DeviceRegion;
std::shared_ptr<DeviceRegion> device_region_create(int device, size_t bytes, std::string& err);
int engine_use_region(Engine* e, std::shared_ptr<DeviceRegion> r, std::string& err);
size_t region_bytes(const DeviceRegion* r);
// f(active rows, all rows) runs on the host after every semi-naive round of
// engine_run (empty f: none).
void engine_on_round(Engine* e, std::function<void(uint64_t, uint64_t)> f);
// on: engine_run leaves the packed labels' ids unwritten until an export or
// checksum needs them; engine_slices reads the packed entries instead.
void engine_defer_unpack(Engine* e, bool on);
int engine_device(Engine* e);
// The share of device memory a batch launch group plans for (tuned).
double tuned_group_mem_share();
gg_stats engine_session_stats(Engine* e, double* alloc_ms);

// The closure of a graph whose edges are already in device memory, computed
// with no device or pinned-host allocation: every buffer comes from a fixed
// device region of noalloc_scratch_bytes() (the current device's own copy).
// The result is written as a CSR (row = (label-1) * n + src, sorted dst ids)
// into the caller's device buffers, on the caller's stream.
size_t noalloc_scratch_bytes();
int engine_run_noalloc(const Grammar& gr, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                       const uint8_t* lab, void* stream, uint64_t* out_ptr, uint32_t* out_col, uint64_t col_cap,
                       uint64_t* total, uint64_t* checksum, uint64_t* per_label, std::string& err);

int device_query(int* count, int preferred, int* selected, char* name, size_t name_cap, uint64_t* total,
                 uint64_t* free_bytes, std::string& err);

}  // namespace gg
