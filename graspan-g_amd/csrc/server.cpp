// Serving loop (gg_server_*; no counterpart in the reference, which closes
// one graph per process): continuous batching at fixpoint-round granularity.
//
// Every slot owns a context (device engine and HIP stream) and runs one
// session at a time: a semi-naive fixpoint that requests join at round
// boundaries (engine_session_*). Before each round the slot admits queued
// requests in arrival order while their estimated footprint fits the device
// budget left by everything in flight; an admitted request's graph takes the
// next vertex range of the session and the label range of its grammar, and
// its edges enter the fixpoint in that round. After each round, a request
// whose rows got no new edges has reached its fixpoint: its result is read
// from the device and published at once, while the other requests of the
// session go on. A session whose next request does not fit stops admitting
// (it drains, so that its memory comes back soon) and ends when its last
// request is done; the slot then starts a new one. Two slots run their
// sessions on separate streams, so their rounds overlap on the device.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"
#include "formats.h"
#include "graspan_g.h"

struct gg_ctx;
namespace gg {
Engine* ctx_engine(gg_ctx* c, int& status, std::string*& err);
}

namespace {

using GrammarKey = std::pair<int, std::vector<int32_t>>;  // (labels, rule triples)

struct Request {
  int64_t id = 0;
  GrammarKey key;
  uint32_t n = 0;
  std::vector<uint32_t> src, dst;
  std::vector<uint8_t> lab;
  uint64_t est = 0;
  bool retried = false;  // re-queued once after its session ran out of device memory
  gg_server_result res{};
  std::vector<uint64_t> counts;
  bool done = false;
};

// At most this many grammars share one session (their labels side by side).
constexpr size_t kSessionGrammars = 8;

struct Live {
  Request* r;
  gg::SessionSlice slice;
};

struct Session {
  bool active = false;
  bool draining = false;
  std::map<GrammarKey, uint32_t> base;  // grammar -> first combined label (0-based)
  uint32_t labels = 0, n = 0;
  int slots = 0;                         // pair-list slots of the combined grammar
  uint64_t est = 0;                      // footprint estimates of every request admitted
  std::vector<Live> live;
  uint32_t rounds = 0, admitted = 0;
  double t0 = 0;
  double t_admit = 0, t_round = 0, t_results = 0;  // host time in each engine call
};

}  // namespace

struct gg_server {
  gg_server_options opt{};
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::unordered_map<int64_t, Request*> all;
  std::vector<GrammarKey> grammars;  // every grammar submitted, in order of first appearance
  int64_t next_id = 1;
  uint64_t in_flight = 0;  // footprint estimates of the requests in running sessions
  uint64_t epoch = 0;      // bumped by every submit and every release of in-flight memory
  std::shared_ptr<gg::DeviceRegion> region;  // device memory of every slot's engine
  bool stopping = false;
  std::vector<std::thread> workers;

  double now() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }

  // GG_SERVER_TRACE=1: one stderr line per session (requests, rounds, time)
  const bool trace = [] {
    const char* t = std::getenv("GG_SERVER_TRACE");
    return t && t[0] == '1';
  }();

  // The session's grammar: the head request's first, then the other known
  // grammars while the engine's limits allow (so later requests of those
  // grammars can join).
  static bool fits_grammar(const Session& s, const GrammarKey& k, int& slots_out) {
    gg::RuleKinds have{}, add = gg::rule_kinds(k.second.data(), int(k.second.size() / 3));
    for (auto& kv : s.base) {
      const gg::RuleKinds r = gg::rule_kinds(kv.first.second.data(), int(kv.first.second.size() / 3));
      have.unary += r.unary;
      have.binary += r.binary;
      have.slots += r.slots;
    }
    slots_out = have.slots + add.slots;
    return s.base.size() < kSessionGrammars && s.labels + uint32_t(k.first) <= uint32_t(gg::kMaxLabels) &&
           have.unary + add.unary <= gg::kMaxRulesPerKind && have.binary + add.binary <= gg::kMaxRulesPerKind;
  }

  // Starts a session on the slot's engine.
  int begin_session(gg::Engine* e, Session& s, const GrammarKey& head, const std::vector<GrammarKey>& known,
                    std::string& err) {
    s = Session{};
    std::vector<GrammarKey> keys{head};
    for (auto& k : known)
      if (k != head) keys.push_back(k);
    for (auto& k : keys) {
      int sl = 0;
      if (!fits_grammar(s, k, sl)) continue;
      s.base[k] = s.labels;
      s.labels += uint32_t(k.first);
      s.slots = sl;
    }
    std::vector<int32_t> rules;
    for (auto& kv : s.base) {
      const auto& r = kv.first.second;
      for (size_t i = 0; i + 2 < r.size(); i += 3) {
        if (r[i + 1] == 0) continue;  // eps: explicit loop edges
        rules.push_back(r[i] + int32_t(kv.second));
        rules.push_back(r[i + 1] + int32_t(kv.second));
        rules.push_back(r[i + 2] ? r[i + 2] + int32_t(kv.second) : 0);
      }
    }
    gg::Grammar cg;
    int st = gg::grammar_from_rules(int(rules.size() / 3), rules.data(), int(std::max(s.labels, 1u)), cg, err);
    if (st == GG_OK) st = gg::engine_session_begin(e, cg, err);
    s.active = st == GG_OK;
    s.t0 = now();
    return st;
  }

  void publish(Request* r, int status, uint64_t total, uint64_t sum, std::vector<uint64_t>&& counts, double start,
               uint32_t cohort, int slot) {
    r->res.status = status;
    r->res.total_edges = total;
    r->res.checksum = sum;
    r->res.start_ms = start;
    r->res.done_ms = now();
    r->res.batch_size = cohort;
    r->res.slot = slot;
    r->counts = std::move(counts);
    r->done = true;
    std::vector<uint32_t>().swap(r->src);
    std::vector<uint32_t>().swap(r->dst);
    std::vector<uint8_t>().swap(r->lab);
  }

  void worker(int slot) {
    gg_options o{opt.device, 0, 0};
    gg_ctx* ctx = nullptr;
    int cs = gg_ctx_create(&o, &ctx);
    std::string* cerr = nullptr;
    std::string err;
    gg::Engine* e = cs == GG_OK ? gg::ctx_engine(ctx, cs, cerr) : nullptr;
    if (e && trace) gg::engine_session_timing(e, true);
    Session s;
    std::unordered_map<Request*, std::pair<double, uint32_t>> start;  // admission time, cohort size
    const uint64_t share = opt.budget_bytes / uint64_t(std::max(1, opt.slots));
    // The whole budget is one device region, allocated once and shared by
    // the slots: sessions grow their buffers every round, and per-buffer
    // hipMalloc and hipFree of tens of GB cost seconds per session.
    if (e) {
      std::shared_ptr<gg::DeviceRegion> r;
      {
        std::lock_guard<std::mutex> lk(mu);
        if (!region) region = gg::device_region_create(gg::engine_device(e), opt.budget_bytes, err);
        r = region;
      }
      cs = r ? gg::engine_use_region(e, r, err) : GG_E_NOMEM;
      if (cs != GG_OK) e = nullptr;
    }
    uint64_t seen = ~0ull;  // epoch at which this slot last found nothing to admit
    auto end_session = [&](bool trim) {
      if (s.active) {
        if (trace) {
          double alloc_ms = 0;
          const gg_stats st = gg::engine_session_stats(e, &alloc_ms);
          std::fprintf(stderr,
                       "[gg-server] slot %d: session of %u requests, %u vertices, %u labels, %u rounds, %.0f ms "
                       "(admit %.0f, rounds %.0f [expand %.0f merge %.0f compact %.0f other %.0f], results %.0f); "
                       "%.1fe9 new edges, device peak %.1f GB, allocation host time so far %.0f ms\n",
                       slot, s.admitted, s.n, s.labels, s.rounds, now() - s.t0, s.t_admit, s.t_round, st.expand_ms,
                       st.merge_ms, st.compact_ms, st.other_ms, s.t_results, double(st.new_edges) / 1e9,
                       double(st.device_bytes_peak) / 1e9, alloc_ms);
        }
        gg::engine_session_end(e, trim || gg::engine_device_bytes(e) > share);
      }
      std::lock_guard<std::mutex> lk(mu);
      if (s.est) {
        in_flight -= s.est;
        ++epoch;  // memory came back: blocked slots look at the queue again
        cv_work.notify_all();
      }
      s = Session{};
    };
    for (;;) {
      std::vector<Request*> take;
      GrammarKey head;
      std::vector<GrammarKey> known;
      bool begin = false;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] {
          return !s.live.empty() || (!queue.empty() && epoch != seen) || (stopping && queue.empty());
        });
        if (s.live.empty() && queue.empty() && stopping) break;
        if (!e) {
          // no device: every queued request fails with the slot's status
          while (!queue.empty()) {
            Request* r = queue.front();
            queue.pop_front();
            publish(r, cs, 0, 0, {}, now(), 1, slot);
          }
          cv_done.notify_all();
          continue;
        }
        if (!s.active && !queue.empty()) {
          head = queue.front()->key;
          known = grammars;
          begin = true;
        }
      }
      if (begin && begin_session(e, s, head, known, err) != GG_OK) {
        std::lock_guard<std::mutex> lk(mu);
        for (auto it = queue.begin(); it != queue.end();)
          if ((*it)->key == head) {
            publish(*it, GG_E_UNSUPPORTED, 0, 0, {}, now(), 1, slot);
            it = queue.erase(it);
          } else {
            ++it;
          }
        cv_done.notify_all();
        s = Session{};
        continue;
      }
      {
        // admission at this round boundary, in arrival order
        std::lock_guard<std::mutex> lk(mu);
        uint64_t nv = s.n;
        for (auto it = queue.begin(); s.active && !s.draining && it != queue.end();) {
          Request* r = *it;
          if (!s.base.count(r->key)) {
            ++it;  // another slot (or a later session) takes it
            continue;
          }
          if (opt.max_batch_requests && take.size() >= opt.max_batch_requests) break;
          // A session holds at most its slot's share of the budget, so that
          // one slot admits while the other drains; a request larger than a
          // share runs alone in its session, and one larger than the budget
          // only when nothing else is on the device.
          const bool alone = s.live.empty() && take.empty() && s.est == 0;
          const bool rows_ok = uint64_t(s.labels) * (nv + r->n) < 0xFFFFFFFFull &&
                               uint64_t(s.slots) * (nv + r->n) < 0xFFFFFFFFull;
          const bool mem_ok = (s.est + r->est <= share || alone) &&
                              (in_flight + r->est <= opt.budget_bytes || in_flight == 0);
          if (r->retried && !alone) break;  // a request that ran out of memory once runs alone
          if (!rows_ok && s.live.empty() && take.empty()) {
            // too many rows for the engine even alone in a session
            publish(r, GG_E_UNSUPPORTED, 0, 0, {}, now(), 1, slot);
            cv_done.notify_all();
            it = queue.erase(it);
            continue;
          }
          if (!rows_ok || !mem_ok) {
            if (!s.live.empty() || !take.empty()) s.draining = true;
            break;
          }
          take.push_back(r);
          nv += r->n;
          in_flight += r->est;
          s.est += r->est;
          it = queue.erase(it);
          if (r->retried) s.draining = true;
        }
        if (take.empty() && s.live.empty()) seen = epoch;  // nothing to do here until something changes
      }
      if (take.empty() && s.live.empty()) {
        end_session(false);
        continue;
      }
      int st = GG_OK;
      if (!take.empty()) {
        // the new graphs' edges in session numbering; eps rules as loops
        uint64_t m = 0;
        for (Request* r : take) {
          m += r->src.size();
          for (size_t i = 0; i + 2 < r->key.second.size(); i += 3)
            if (r->key.second[i + 1] == 0) m += r->n;
        }
        std::vector<uint32_t> src(m), dst(m);
        std::vector<uint8_t> lab(m);
        uint64_t p = 0;
        uint32_t v0 = s.n;
        const double t_adm = now();
        for (Request* r : take) {
          const uint32_t lb = s.base[r->key];
          for (size_t k = 0; k < r->src.size(); ++k) {
            src[p] = r->src[k] + v0;
            dst[p] = r->dst[k] + v0;
            lab[p++] = uint8_t(r->lab[k] + lb);
          }
          for (size_t i = 0; i + 2 < r->key.second.size(); i += 3) {
            if (r->key.second[i + 1] != 0) continue;
            const uint8_t el = uint8_t(r->key.second[i] + int32_t(lb));
            for (uint32_t v = 0; v < r->n; ++v) {
              src[p] = dst[p] = v0 + v;
              lab[p++] = el;
            }
          }
          s.live.push_back({r, {v0, r->n, lb, uint32_t(r->key.first)}});
          start[r] = {t_adm, uint32_t(take.size())};
          v0 += r->n;
        }
        const double a0 = now();
        st = gg::engine_session_admit(e, v0, src.data(), dst.data(), lab.data(), m, err);
        s.t_admit += now() - a0;
        s.n = v0;
        s.admitted += uint32_t(take.size());
      }
      std::vector<uint64_t> delta;
      if (st == GG_OK) {
        std::vector<gg::SessionSlice> sl;
        for (auto& l : s.live) sl.push_back(l.slice);
        const double a0 = now();
        st = gg::engine_session_round(e, sl, delta, err);
        s.t_round += now() - a0;
        ++s.rounds;
      }
      if (st != GG_OK) {
        // The session failed (most likely out of device memory): its
        // unfinished requests go back to the front of the queue once, to run
        // alone; a request that fails again reports the error.
        {
          std::lock_guard<std::mutex> lk(mu);
          std::vector<Request*> back;
          for (auto& l : s.live) {
            if (!l.r->retried && st == GG_E_NOMEM) {
              l.r->retried = true;
              back.push_back(l.r);
            } else {
              publish(l.r, st, 0, 0, {}, start[l.r].first, start[l.r].second, slot);
            }
            start.erase(l.r);
          }
          for (auto it = back.rbegin(); it != back.rend(); ++it) queue.push_front(*it);
          ++epoch;
          if (trace) std::fprintf(stderr, "[gg-server] slot %d: session failed (%s)\n", slot, err.c_str());
        }
        cv_done.notify_all();
        s.live.clear();
        end_session(true);
        continue;
      }
      // requests whose rows got no new edges are done
      std::vector<Live> done, rest;
      for (size_t i = 0; i < s.live.size(); ++i) (delta[i] == 0 ? done : rest).push_back(s.live[i]);
      if (!done.empty()) {
        std::vector<gg::SessionSlice> sl;
        for (auto& l : done) sl.push_back(l.slice);
        std::vector<uint64_t> counts, sums;
        const double a0 = now();
        st = gg::engine_session_results(e, sl, counts, sums, err);
        s.t_results += now() - a0;
        {
          std::lock_guard<std::mutex> lk(mu);
          size_t c = 0;
          for (size_t q = 0; q < done.size(); ++q) {
            Request* r = done[q].r;
            std::vector<uint64_t> per;
            uint64_t tot = 0;
            if (st == GG_OK) {
              per.assign(counts.begin() + long(c), counts.begin() + long(c + done[q].slice.lg));
              for (uint64_t x : per) tot += x;
            }
            c += done[q].slice.lg;
            publish(r, st, tot, st == GG_OK ? sums[q] : 0, std::move(per), start[r].first, start[r].second, slot);
            start.erase(r);
          }
        }
        cv_done.notify_all();
      }
      s.live.swap(rest);
      // a session that holds more than half its region takes no one else:
      // its live graphs still grow, and finished graphs keep their memory
      // until the session ends
      if (!s.draining && gg::engine_device_bytes(e) > share / 2) s.draining = true;
      if (s.live.empty()) {
        // nothing left in the session: it ends unless queued requests can
        // still join it (its finished graphs' memory stays until it does)
        bool more = false;
        {
          std::lock_guard<std::mutex> lk(mu);
          if (!s.draining)
            for (Request* r : queue)
              if (s.base.count(r->key) && !r->retried && s.est + r->est <= share &&
                  in_flight + r->est <= opt.budget_bytes) {
                more = true;
                break;
              }
        }
        if (!more) end_session(false);
      }
    }
    if (ctx) gg_ctx_destroy(ctx);
  }
};

extern "C" {

int gg_server_create(const gg_server_options* opts, gg_server** out) {
  if (!opts || !out || opts->slots < 1 || opts->budget_bytes == 0) return GG_E_ARG;
  auto* s = new gg_server();
  s->opt = *opts;
  s->t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < opts->slots; ++i) s->workers.emplace_back(&gg_server::worker, s, i);
  *out = s;
  return GG_OK;
}

int gg_server_submit(gg_server* s, int num_rules, const int32_t* rules, int num_labels, uint32_t n, uint64_t m,
                     const uint32_t* src, const uint32_t* dst, const uint8_t* lab, int64_t* id) {
  if (!s || num_rules < 0 || (num_rules && !rules) || num_labels < 1 || num_labels > gg::kMaxLabels || n == 0 ||
      (m && (!src || !dst || !lab)))
    return GG_E_ARG;
  gg::Grammar tmp;
  std::string err;
  if (gg::grammar_from_rules(num_rules, rules, num_labels, tmp, err) != GG_OK) return GG_E_ARG;
  for (uint64_t i = 0; i < m; ++i)
    if (src[i] >= n || dst[i] >= n || lab[i] < 1 || lab[i] > num_labels) return GG_E_ARG;
  auto* r = new Request();
  // requests are closed to their least fixpoint: their rules as planned for
  // it (eps rules stay, and become loop edges when the request is admitted)
  r->key = {num_labels, gg::plan_fixpoint_rules(rules, num_rules, gg::kMaxRulesPerKind)};
  r->n = n;
  r->src.assign(src, src + m);
  r->dst.assign(dst, dst + m);
  r->lab.assign(lab, lab + m);
  r->est = s->opt.bytes_per_vertex * n + 64 * m;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) {
      delete r;
      return GG_E_STATE;
    }
    if (std::find(s->grammars.begin(), s->grammars.end(), r->key) == s->grammars.end())
      s->grammars.push_back(r->key);
    r->id = s->next_id++;
    r->res.submit_ms = s->now();
    s->all[r->id] = r;
    s->queue.push_back(r);
    ++s->epoch;
    if (id) *id = r->id;
  }
  s->cv_work.notify_all();
  return GG_OK;
}

int gg_server_wait(gg_server* s, int64_t id, gg_server_result* res, uint64_t* per_label) {
  if (!s) return GG_E_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) return GG_E_ARG;
  Request* r = it->second;
  s->cv_done.wait(lk, [&] { return r->done; });
  if (res) *res = r->res;
  if (per_label)
    for (size_t l = 0; l < r->counts.size(); ++l) per_label[l] = r->counts[l];
  return GG_OK;
}

double gg_server_now_ms(const gg_server* s) { return s ? s->now() : 0.0; }

int gg_server_destroy(gg_server* s) {
  if (!s) return GG_E_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->workers) t.join();
  for (auto& kv : s->all) delete kv.second;
  delete s;
  return GG_OK;
}

}  // extern "C"
