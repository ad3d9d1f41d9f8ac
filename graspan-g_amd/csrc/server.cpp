// Serving loop over the batched closure (gg_server_*; no counterpart in the
// reference). See include/graspan_g.h for the admission policy.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "graspan_g.h"

namespace {

struct Request {
  int64_t id = 0;
  std::vector<int32_t> rules;
  int L = 0;
  uint32_t n = 0;
  std::vector<uint32_t> src, dst;
  std::vector<uint8_t> lab;
  uint64_t est = 0;
  gg_server_result res{};
  std::vector<uint64_t> counts;
  bool done = false;
};

}  // namespace

struct gg_server {
  gg_server_options opt{};
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::unordered_map<int64_t, Request*> all;
  int64_t next_id = 1;
  uint64_t in_flight = 0;
  bool stopping = false;
  std::vector<std::thread> workers;

  double now() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }

  // GG_SERVER_TRACE=1: one line per batch on stderr (sizes and phase times)
  const bool trace = [] {
    const char* t = std::getenv("GG_SERVER_TRACE");
    return t && t[0] == '1';
  }();

  void worker(int slot) {
    gg_options o{opt.device, 0, 0};
    gg_ctx* ctx = nullptr;
    const int cs = gg_ctx_create(&o, &ctx);
    for (;;) {
      std::vector<Request*> take;
      uint64_t est = 0;
      {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
          if (!queue.empty()) {
            // admit in arrival order while the footprint fits the budget left
            // by the other slots; a request larger than the whole budget is
            // admitted alone when nothing else is running
            while (!queue.empty() && (opt.max_batch_requests == 0 || take.size() < opt.max_batch_requests)) {
              Request* r = queue.front();
              const bool fits = in_flight + est + r->est <= opt.budget_bytes;
              const bool alone = take.empty() && in_flight == 0;
              if (!fits && !alone) break;
              take.push_back(r);
              est += r->est;
              queue.pop_front();
            }
            if (!take.empty()) {
              in_flight += est;
              break;
            }
          }
          if (stopping && queue.empty()) {
            if (ctx) gg_ctx_destroy(ctx);
            return;
          }
          cv_work.wait(lk);
        }
      }
      const double start = now();
      // one batch on the slot's context; a batch that runs out of device
      // memory is split in halves and retried (the footprint estimate is
      // only an estimate), down to single requests
      std::vector<std::vector<uint64_t>> counts(take.size());
      std::vector<uint64_t> totals(take.size(), 0), sums(take.size(), 0);
      std::vector<int> rstatus(take.size(), cs);
      double t_add = 0, t_run = 0;
      gg_stats bst{};
      std::function<void(size_t, size_t)> exec = [&](size_t lo, size_t hi) {
        int status = cs;
        gg_batch* b = nullptr;
        std::vector<int> ids(hi - lo, -1);
        const double a0 = now();
        if (status == GG_OK) status = gg_batch_create(ctx, &b);
        for (size_t i = lo; status == GG_OK && i < hi; ++i) {
          Request* r = take[i];
          status = gg_batch_add(b, int(r->rules.size() / 3), r->rules.data(), r->L, r->n, r->src.size(),
                                r->src.data(), r->dst.data(), r->lab.data(), &ids[i - lo]);
        }
        const double a1 = now();
        if (status == GG_OK) status = gg_batch_run(b, 0, 0, &bst);
        t_add += a1 - a0;
        t_run += now() - a1;
        for (size_t i = lo; status == GG_OK && i < hi; ++i) {
          counts[i].assign(size_t(take[i]->L), 0);
          status = gg_batch_result(b, ids[i - lo], counts[i].data(), &totals[i], &sums[i]);
        }
        if (b) gg_batch_destroy(b);
        // A slot keeps its device memory for its next batch (freeing and
        // re-allocating tens of GB per batch costs seconds) while the batch
        // stayed within the slot's share of the budget; otherwise, or after
        // an error, the memory goes back for the other slots.
        const uint64_t share = opt.budget_bytes / uint64_t(opt.slots);
        if (ctx && (status != GG_OK || bst.device_bytes_peak > share)) gg_closure_release(ctx);
        if (status == GG_E_NOMEM && hi - lo > 1) {
          const size_t mid = lo + (hi - lo) / 2;
          exec(lo, mid);
          exec(mid, hi);
          return;
        }
        for (size_t i = lo; i < hi; ++i) rstatus[i] = status;
      };
      exec(0, take.size());
      const double t_res = now();
      const double end = now();
      if (trace) {
        uint64_t nv = 0, edges = 0;
        for (size_t i = 0; i < take.size(); ++i) {
          nv += take[i]->n;
          edges += totals[i];
        }
        std::fprintf(stderr,
                     "[gg-server] slot %d: %zu requests, %llu vertices, %llu closure edges: add %.0f ms, run %.0f ms "
                     "(last fixpoint %.0f, %u rounds), total %.0f ms\n",
                     slot, take.size(), (unsigned long long)nv, (unsigned long long)edges, t_add, t_run,
                     bst.fixpoint_ms, bst.rounds, t_res - start);
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        for (size_t i = 0; i < take.size(); ++i) {
          Request* r = take[i];
          r->res.status = rstatus[i];
          r->res.total_edges = totals[i];
          r->res.checksum = sums[i];
          r->res.start_ms = start;
          r->res.done_ms = end;
          r->res.batch_size = uint32_t(take.size());
          r->res.slot = slot;
          r->counts = std::move(counts[i]);
          r->done = true;
          // inputs are no longer needed
          std::vector<uint32_t>().swap(r->src);
          std::vector<uint32_t>().swap(r->dst);
          std::vector<uint8_t>().swap(r->lab);
        }
        in_flight -= est;
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
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
  if (!s || num_rules < 0 || (num_rules && !rules) || num_labels < 1 || n == 0 || (m && (!src || !dst || !lab)))
    return GG_E_ARG;
  auto* r = new Request();
  r->rules.assign(rules, rules + 3 * num_rules);
  r->L = num_labels;
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
    r->id = s->next_id++;
    r->res.submit_ms = s->now();
    s->all[r->id] = r;
    s->queue.push_back(r);
    if (id) *id = r->id;
  }
  s->cv_work.notify_one();
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
