// Continuous-batching proof server: request queue, HBM-budget admission and
// the scheduler thread that runs one step at a time (server part of the C ABI).
//
// Each step: admit queued requests in arrival order while their resident
// state fits the budget, launch the next round of every admitted request
// (one batched launch per request kind), download the outputs of the requests
// that complete with this step, wait for the step, and hand those results to
// pollers. A request therefore joins the batch at the first step after it
// arrives, whatever the rounds the others are at.
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/zorch_amd.h"
#include "runtime.h"

namespace zs {

using Clock = std::chrono::steady_clock;
using KindKey = std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, bool>;

struct Pending {
  uint64_t id = 0;
  zs_prove_args a{};
  zs_transcript t{};
  std::vector<uint8_t> claim, eq_point;
  Clock::time_point t_submit;
  ServeKind* kind = nullptr;
  uint64_t bytes = 0;
};

struct Finished {
  int rc = ZS_OK;
  std::string err;
  ServeResult res;
  double latency = 0.0;
};

struct Active {
  uint64_t id = 0;
  ServeKind* kind = nullptr;
  std::unique_ptr<ServeReq> req;
  Clock::time_point t_submit;
};

}  // namespace zs

using namespace zs;

struct zs_server {
  zs_ctx* ctx = nullptr;
  uint64_t budget = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Pending> queue;
  std::map<uint64_t, Finished> finished;
  std::deque<uint64_t> completed_ids;
  std::set<uint64_t> known;  // submitted, not yet polled
  std::map<KindKey, std::unique_ptr<ServeKind>> kinds;
  bool stop = false;
  uint64_t next_id = 1;
  uint64_t resident = 0;
  zs_server_stats stats{};
  std::thread th;
  // Scheduler thread only.
  std::vector<Active> active;
};

namespace {

// Record results (or failures) for pollers.
void publish(zs_server* S, std::vector<std::pair<uint64_t, Finished>>& done, uint64_t freed) {
  std::lock_guard<std::mutex> lk(S->mu);
  S->resident -= freed;
  for (auto& d : done) {
    if (d.second.rc == ZS_OK) ++S->stats.completed; else ++S->stats.failed;
    S->completed_ids.push_back(d.first);
    S->finished[d.first] = std::move(d.second);
  }
  S->stats.active = S->active.size();
  S->stats.resident_bytes = S->resident;
  done.clear();
  S->cv_done.notify_all();
}

void fail_all_active(zs_server* S, int rc, const std::string& err) {
  std::vector<std::pair<uint64_t, Finished>> done;
  uint64_t freed = 0;
  for (Active& a : S->active) {
    a.kind->release(S->ctx, a.req.get());
    freed += a.req->bytes;
    Finished f;
    f.rc = rc;
    f.err = err;
    done.emplace_back(a.id, std::move(f));
  }
  S->active.clear();
  publish(S, done, freed);
}

void server_loop(zs_server* S) {
  zs_ctx* ctx = S->ctx;
  (void)hipSetDevice(ctx->device);
  std::vector<Pending> admit;
  std::vector<std::pair<uint64_t, Finished>> done;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(S->mu);
      if (S->active.empty())
        S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty(); });
      if (S->stop) break;
      // Arrival order; a request that can never fit fails instead of blocking.
      while (!S->queue.empty()) {
        Pending& p = S->queue.front();
        if (S->resident + p.bytes > S->budget) {
          if (!S->active.empty() || !admit.empty()) break;
          Finished f;
          f.rc = ZS_EBUDGET;
          f.err = "request needs " + std::to_string(p.bytes) + " device bytes; server budget is " +
                  std::to_string(S->budget);
          ++S->stats.failed;
          S->completed_ids.push_back(p.id);
          S->finished[p.id] = std::move(f);
          S->queue.pop_front();
          S->cv_done.notify_all();
          continue;
        }
        S->resident += p.bytes;
        admit.push_back(std::move(p));
        S->queue.pop_front();
      }
      S->stats.queued = S->queue.size();
      S->stats.resident_bytes = S->resident;
      if (S->resident > S->stats.peak_resident_bytes) S->stats.peak_resident_bytes = S->resident;
    }
    const auto t_step = Clock::now();
    uint64_t freed = 0;
    for (Pending& p : admit) {
      p.a.claim = p.claim.data();
      p.a.eq_point = p.eq_point.empty() ? nullptr : p.eq_point.data();
      ServeReq* r = nullptr;
      const int rc = p.kind->admit(ctx, p.a, p.t, &r);
      if (rc) {
        Finished f;
        f.rc = rc;
        f.err = zs_last_error();
        freed += p.bytes;
        done.emplace_back(p.id, std::move(f));
        continue;
      }
      Active a;
      a.id = p.id;
      a.kind = p.kind;
      a.req.reset(r);
      a.t_submit = p.t_submit;
      S->active.push_back(std::move(a));
    }
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->stats.admitted += admit.size();
      if (S->active.size() > S->stats.peak_active) S->stats.peak_active = S->active.size();
    }
    admit.clear();
    if (!done.empty() || freed) publish(S, done, freed);
    if (S->active.empty()) continue;
    // One step: the next round of every admitted request, one launch per kind.
    std::map<ServeKind*, std::vector<ServeReq*>> groups;
    for (Active& a : S->active) groups[a.kind].push_back(a.req.get());
    int rc = ZS_OK;
    for (auto& g : groups)
      if ((rc = g.first->step(ctx, g.second, ctx->stream))) break;
    if (!rc) {
      for (Active& a : S->active)
        if (a.req->finishing && (rc = a.kind->download(a.req.get(), ctx->stream))) break;
    }
    if (!rc) {
      const hipError_t e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) rc = hip_fail(e, "hipStreamSynchronize");
    }
    if (rc) {
      fail_all_active(S, rc, zs_last_error());
      continue;
    }
    const auto now = Clock::now();
    std::vector<Active> still;
    for (Active& a : S->active) {
      if (!a.req->finishing) {
        still.push_back(std::move(a));
        continue;
      }
      Finished f;
      a.kind->collect(a.req.get(), &f.res);
      f.latency = std::chrono::duration<double>(now - a.t_submit).count();
      a.kind->release(ctx, a.req.get());
      freed += a.req->bytes;
      done.emplace_back(a.id, std::move(f));
    }
    S->active.swap(still);
    {
      std::lock_guard<std::mutex> lk(S->mu);
      ++S->stats.steps;
      S->stats.busy_s += std::chrono::duration<double>(now - t_step).count();
    }
    publish(S, done, freed);
  }
  for (Active& a : S->active) a.kind->release(ctx, a.req.get());
  S->active.clear();
}

int take_result(zs_server* S, std::unique_lock<std::mutex>& lk, uint64_t id, zs_proof* out,
                zs_transcript* t, double* latency, int* done) {
  auto it = S->finished.find(id);
  if (it == S->finished.end()) {
    if (!S->known.count(id)) return fail(ZS_EINVAL, "unknown request id");
    *done = 0;
    return ZS_OK;
  }
  Finished f = std::move(it->second);
  S->finished.erase(it);
  S->known.erase(id);
  lk.unlock();
  *done = 1;
  if (f.rc) return fail(f.rc, f.err);
  if (out) {
    if (out->msgs) memcpy(out->msgs, f.res.msgs.data(), f.res.msgs.size());
    if (out->point) memcpy(out->point, f.res.point.data(), f.res.point.size());
    if (out->value) memcpy(out->value, f.res.value.data(), f.res.value.size());
    if (out->final_tables)
      memcpy(out->final_tables, f.res.final_tables.data(), f.res.final_tables.size());
  }
  if (t) memcpy(&t->sha, f.res.tr_state.data(), f.res.tr_state.size());
  if (latency) *latency = f.latency;
  return ZS_OK;
}

}  // namespace

extern "C" {

int zs_server_create(zs_ctx* ctx, uint64_t budget, zs_server** out) {
  if (!ctx || !out) return fail(ZS_EINVAL, "null argument");
  auto S = std::make_unique<zs_server>();
  S->ctx = ctx;
  S->budget = budget ? budget : ctx->pool->budget();
  if (S->budget > ctx->pool->budget()) S->budget = ctx->pool->budget();
  S->stats.budget_bytes = S->budget;
  zs_server* raw = S.get();
  raw->th = std::thread(server_loop, raw);
  *out = S.release();
  return ZS_OK;
}

int zs_server_submit(zs_server* S, const zs_prove_args* args, const zs_transcript* t, uint64_t* id) {
  if (!S || !args || !t || !id) return fail(ZS_EINVAL, "null argument");
  uint32_t npts = 0;
  int rc = validate_args(*args, t, &npts);
  if (rc) return rc;
  const bool eq = args->mode == ZS_MODE_EQ;
  const uint32_t cb = field_bytes(args->chal_field);
  Pending p;
  p.a = *args;
  p.t = *t;
  p.claim.assign((const uint8_t*)args->claim, (const uint8_t*)args->claim + cb);
  if (eq) p.eq_point.assign((const uint8_t*)args->eq_point,
                            (const uint8_t*)args->eq_point + (size_t)args->rounds * cb);
  p.t_submit = Clock::now();
  const KindKey key{args->table_field, args->chal_field, args->num_tables, npts, eq};
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->stop) return fail(ZS_EINVAL, "server is stopping");
  auto& kind = S->kinds[key];
  if (!kind) {
    kind.reset(make_serve_kind(args->table_field, args->chal_field, args->num_tables, npts, eq));
    if (!kind) return fail(ZS_EINVAL, "unsupported field / degree combination");
  }
  p.kind = kind.get();
  p.bytes = kind->bytes_needed(*args);
  p.id = S->next_id++;
  *id = p.id;
  S->known.insert(p.id);
  S->queue.push_back(std::move(p));
  ++S->stats.submitted;
  S->stats.queued = S->queue.size();
  S->cv_work.notify_one();
  return ZS_OK;
}

int zs_server_poll(zs_server* S, uint64_t id, zs_proof* out, zs_transcript* t, double* latency,
                   int* done) {
  if (!S || !done) return fail(ZS_EINVAL, "null argument");
  std::unique_lock<std::mutex> lk(S->mu);
  return take_result(S, lk, id, out, t, latency, done);
}

int zs_server_wait(zs_server* S, uint64_t id, zs_proof* out, zs_transcript* t, double* latency,
                   double timeout_s, int* done) {
  if (!S || !done) return fail(ZS_EINVAL, "null argument");
  std::unique_lock<std::mutex> lk(S->mu);
  S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s),
                      [&] { return S->finished.count(id) || !S->known.count(id); });
  return take_result(S, lk, id, out, t, latency, done);
}

int zs_server_completed(zs_server* S, uint64_t* ids, size_t cap, size_t* count) {
  if (!S || !count || (cap && !ids)) return fail(ZS_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  size_t n = 0;
  while (n < cap && !S->completed_ids.empty()) {
    ids[n++] = S->completed_ids.front();
    S->completed_ids.pop_front();
  }
  *count = n;
  return ZS_OK;
}

int zs_server_stats_read(zs_server* S, zs_server_stats* out) {
  if (!S || !out) return fail(ZS_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  *out = S->stats;
  out->queued = S->queue.size();
  out->resident_bytes = S->resident;
  return ZS_OK;
}

void zs_server_destroy(zs_server* S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->th.joinable()) S->th.join();
  delete S;
}

}  // extern "C"
