// Continuous-batching proof server: request queue, HBM-budget admission and
// the scheduler thread that runs one step at a time (server part of the C ABI).
//
// Each step: admit queued requests in arrival order while their resident
// state fits the budget, launch the next round of every admitted request
// (one batched launch per request kind), download the outputs of the requests
// that complete with this step and record an event. Up to STEPS_IN_FLIGHT
// steps are queued on the stream at once: the host builds the next step
// while the device runs the previous ones (the rounds a step launches are
// fixed on the host when it is built), and the results of a step are handed
// to pollers once its event has completed. A request therefore joins the
// batch at the first step built after it arrives, whatever the rounds the
// others are at.
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
  std::vector<uint8_t> claim, eq_point, eq_mass;
  bool has_wire = false;
  zs_wire wire{};              // copied wire form; its nodes in `nodes`
  std::vector<uint8_t> nodes;
  Clock::time_point t_submit;
  ServeKind* kind = nullptr;
  uint64_t bytes = 0;
  uint32_t npts = 0, mlen = 0;
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

// A step queued on the stream: its completion event and the requests that
// complete with it (their outputs are being downloaded).
struct Flight {
  hipEvent_t ev = nullptr;
  std::vector<Active> done;
  Clock::time_point t_launch;
};
constexpr size_t STEPS_IN_FLIGHT = 3;

// A large request proved whole on one of the server's own streams (lanes)
// by the single-proof kernels, so that its streaming rounds overlap the
// latency-bound rounds of the others; chained rounds are off there (their
// blocks wait for one another, which concurrent launches could starve).
struct Direct {
  uint64_t id = 0;
  std::unique_ptr<Instance> inst;
  hipEvent_t ev = nullptr;
  uint64_t bytes = 0;
  Clock::time_point t_submit;
  uint32_t n = 0, npts = 0, mlen = 0, k = 0, cb = 0;
  bool eq = false;
};
constexpr uint32_t DIRECT_MIN_ROUNDS = 20;
// Streams of the direct lanes. A process gets four hardware queues (HIP's
// default); the server's step stream and two lanes leave one for the
// caller's own stream: two streams sharing a queue run one after the other,
// which with three lanes and a caller's stream throttled the served steps.
constexpr int LANES = 2;
constexpr size_t TR_BYTES = sizeof(zs_transcript) - offsetof(zs_transcript, sha);

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
  std::deque<Flight> flights;
  std::vector<hipEvent_t> free_events;
  Clock::time_point busy_until{};  // end of the busy time already counted
  std::vector<hipStream_t> lanes;
  size_t lane_next = 0;
  std::vector<Direct> directs;
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
  S->stats.active = S->active.size() + S->directs.size();
  S->stats.resident_bytes = S->resident;
  done.clear();
  S->cv_done.notify_all();
}

// After a device error: wait for the stream, then fail every admitted and
// in-flight request.
void fail_all_active(zs_server* S, int rc, const std::string& err) {
  (void)hipStreamSynchronize(S->ctx->stream);
  for (hipStream_t l : S->lanes) (void)hipStreamSynchronize(l);
  std::vector<std::pair<uint64_t, Finished>> done;
  uint64_t freed = 0;
  auto fail_one = [&](Active& a) {
    a.kind->release(S->ctx, a.req.get());
    freed += a.req->bytes;
    Finished f;
    f.rc = rc;
    f.err = err;
    done.emplace_back(a.id, std::move(f));
  };
  for (Active& a : S->active) fail_one(a);
  S->active.clear();
  for (Flight& fl : S->flights) {
    for (Active& a : fl.done) fail_one(a);
    S->free_events.push_back(fl.ev);
  }
  S->flights.clear();
  for (Direct& d : S->directs) {
    freed += d.bytes;
    Finished f;
    f.rc = rc;
    f.err = err;
    done.emplace_back(d.id, std::move(f));
    if (d.ev) S->free_events.push_back(d.ev);
  }
  S->directs.clear();
  publish(S, done, freed);
}

// Start request p whole on the next lane; a failure is reported in `done`.
void direct_start(zs_server* S, Pending& p, std::vector<std::pair<uint64_t, Finished>>& done,
                  uint64_t& freed) {
  zs_ctx* ctx = S->ctx;
  const bool eq = p.a.mode == ZS_MODE_EQ;
  Direct d;
  d.id = p.id;
  d.bytes = p.bytes;
  d.t_submit = p.t_submit;
  d.n = p.a.rounds;
  d.npts = p.npts;
  d.mlen = p.mlen;
  d.k = p.a.num_tables;
  d.cb = field_bytes(p.a.chal_field);
  d.eq = eq;
  d.inst.reset(make_instance(p.a.table_field, p.a.chal_field, p.a.num_tables, p.npts, eq));
  int rc = d.inst ? ZS_OK : fail(ZS_EINVAL, "unsupported field / degree combination");
  const hipStream_t lane = S->lanes[S->lane_next++ % S->lanes.size()];
  TuneParams tune = ctx->tune;
  tune.chain_pairs = 0;
  if (!rc) rc = d.inst->setup(ctx, p.a, p.t, lane, tune);
  if (!rc) rc = d.inst->enqueue(0, d.n, lane);
  if (!rc) rc = d.inst->download(lane);
  if (!rc) {
    if (S->free_events.empty()) {
      hipEvent_t e = nullptr;
      const hipError_t he = hipEventCreateWithFlags(&e, hipEventDisableTiming);
      if (he != hipSuccess) rc = hip_fail(he, "hipEventCreateWithFlags");
      else S->free_events.push_back(e);
    }
  }
  if (!rc) {
    d.ev = S->free_events.back();
    S->free_events.pop_back();
    const hipError_t he = hipEventRecord(d.ev, lane);
    if (he != hipSuccess) {
      S->free_events.push_back(d.ev);
      d.ev = nullptr;
      rc = hip_fail(he, "hipEventRecord");
    }
  }
  if (rc) {
    if (d.inst) {
      (void)hipStreamSynchronize(lane);
      d.inst->free_buffers();
    }
    Finished f;
    f.rc = rc;
    f.err = zs_last_error();
    freed += p.bytes;
    done.emplace_back(p.id, std::move(f));
    return;
  }
  S->directs.push_back(std::move(d));
}

// Publish every lane proof whose event has completed.
void collect_directs(zs_server* S) {
  std::vector<std::pair<uint64_t, Finished>> done;
  uint64_t freed = 0;
  std::vector<Direct> still;
  for (Direct& d : S->directs) {
    const hipError_t q = hipEventQuery(d.ev);
    if (q == hipErrorNotReady) {
      still.push_back(std::move(d));
      continue;
    }
    Finished f;
    if (q != hipSuccess) {
      f.rc = hip_fail(q, "hipEventQuery");
      f.err = zs_last_error();
    } else {
      ServeResult& r = f.res;
      r.msgs.resize((size_t)d.n * d.mlen * d.cb);
      r.point.resize((size_t)d.n * d.cb);
      r.value.resize(d.cb);
      r.final_tables.resize((size_t)(d.k + (d.eq ? 1 : 0)) * d.cb);
      zs_proof pr{r.msgs.data(), r.point.data(), r.value.data(), r.final_tables.data()};
      zs_transcript tr{};
      f.rc = d.inst->collect(&pr, &tr);
      if (f.rc) f.err = zs_last_error();
      r.tr_state.assign((const uint8_t*)&tr.sha, (const uint8_t*)&tr.sha + TR_BYTES);
      f.latency = std::chrono::duration<double>(Clock::now() - d.t_submit).count();
    }
    d.inst->free_buffers();
    S->free_events.push_back(d.ev);
    freed += d.bytes;
    done.emplace_back(d.id, std::move(f));
  }
  S->directs.swap(still);
  if (!done.empty() || freed) publish(S, done, freed);
}

// Hand the results of the oldest in-flight step to pollers (its event has
// completed).
void collect_flight(zs_server* S) {
  Flight fl = std::move(S->flights.front());
  S->flights.pop_front();
  S->free_events.push_back(fl.ev);
  const auto now = Clock::now();
  std::vector<std::pair<uint64_t, Finished>> done;
  uint64_t freed = 0;
  for (Active& a : fl.done) {
    Finished f;
    a.kind->collect(a.req.get(), &f.res);
    f.latency = std::chrono::duration<double>(now - a.t_submit).count();
    a.kind->release(S->ctx, a.req.get());
    freed += a.req->bytes;
    done.emplace_back(a.id, std::move(f));
  }
  // Steps in flight overlap: count the union of [launch, collected] spans.
  const auto from = fl.t_launch > S->busy_until ? fl.t_launch : S->busy_until;
  if (now > from) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stats.busy_s += std::chrono::duration<double>(now - from).count();
  }
  if (now > S->busy_until) S->busy_until = now;
  publish(S, done, freed);
}

void server_loop(zs_server* S) {
  zs_ctx* ctx = S->ctx;
  (void)hipSetDevice(ctx->device);
  for (int i = 0; i < LANES; ++i) {
    hipStream_t l = nullptr;
    if (hipStreamCreateWithFlags(&l, hipStreamNonBlocking) != hipSuccess) break;
    S->lanes.push_back(l);
  }
  std::vector<Pending> admit;
  std::vector<std::pair<uint64_t, Finished>> done;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(S->mu);
      auto ready = [&] { return S->stop || !S->queue.empty(); };
      if (S->active.empty() && S->flights.empty()) {
        if (S->directs.empty()) S->cv_work.wait(lk, ready);
        else S->cv_work.wait_for(lk, std::chrono::microseconds(50), ready);
      }
      if (S->stop) break;
      // Arrival order; a request that can never fit fails instead of blocking.
      while (!S->queue.empty()) {
        Pending& p = S->queue.front();
        if (S->resident + p.bytes > S->budget) {
          if (!S->active.empty() || !S->flights.empty() || !S->directs.empty() || !admit.empty())
            break;
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
    uint64_t freed = 0;
    for (Pending& p : admit) {
      p.a.claim = p.claim.data();
      p.a.eq_point = p.eq_point.empty() ? nullptr : p.eq_point.data();
      p.a.eq_mass = p.eq_mass.empty() ? nullptr : p.eq_mass.data();
      p.wire.nodes = p.nodes.empty() ? nullptr : p.nodes.data();
      p.a.wire = p.has_wire ? &p.wire : nullptr;
      if (p.a.rounds >= DIRECT_MIN_ROUNDS && !S->lanes.empty()) {
        direct_start(S, p, done, freed);
        continue;
      }
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
      const size_t live = S->active.size() + S->directs.size();
      if (live > S->stats.peak_active) S->stats.peak_active = live;
    }
    admit.clear();
    if (!done.empty() || freed) publish(S, done, freed);
    if (!S->active.empty() && S->flights.size() < STEPS_IN_FLIGHT) {
      // One step: the next round of every admitted request, one launch per kind.
      Flight fl;
      fl.t_launch = Clock::now();
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
        if (S->free_events.empty()) {
          hipEvent_t e = nullptr;
          const hipError_t he = hipEventCreateWithFlags(&e, hipEventDisableTiming);
          if (he != hipSuccess) rc = hip_fail(he, "hipEventCreateWithFlags");
          else S->free_events.push_back(e);
        }
      }
      if (!rc) {
        fl.ev = S->free_events.back();
        S->free_events.pop_back();
        const hipError_t he = hipEventRecord(fl.ev, ctx->stream);
        if (he != hipSuccess) {
          S->free_events.push_back(fl.ev);
          rc = hip_fail(he, "hipEventRecord");
        }
      }
      if (rc) {
        fail_all_active(S, rc, zs_last_error());
        continue;
      }
      std::vector<Active> still;
      for (Active& a : S->active) {
        if (a.req->finishing) fl.done.push_back(std::move(a));
        else still.push_back(std::move(a));
      }
      S->active.swap(still);
      S->flights.push_back(std::move(fl));
      {
        std::lock_guard<std::mutex> lk(S->mu);
        ++S->stats.steps;
        S->stats.active = S->active.size() + S->directs.size();
      }
    } else if (!S->flights.empty()) {
      // Nothing to launch: wait for the oldest step.
      const hipError_t he = hipEventSynchronize(S->flights.front().ev);
      if (he != hipSuccess) {
        fail_all_active(S, hip_fail(he, "hipEventSynchronize"), zs_last_error());
        continue;
      }
    }
    // Results of every completed step, oldest first, and of finished lanes.
    while (!S->flights.empty()) {
      const hipError_t q = hipEventQuery(S->flights.front().ev);
      if (q == hipErrorNotReady) break;
      if (q != hipSuccess) {
        fail_all_active(S, hip_fail(q, "hipEventQuery"), zs_last_error());
        break;
      }
      collect_flight(S);
    }
    if (!S->directs.empty()) collect_directs(S);
  }
  // Stopping: let queued steps finish, then drop every request.
  (void)hipStreamSynchronize(ctx->stream);
  for (Flight& fl : S->flights) {
    for (Active& a : fl.done) a.kind->release(ctx, a.req.get());
    S->free_events.push_back(fl.ev);
  }
  S->flights.clear();
  for (Active& a : S->active) a.kind->release(ctx, a.req.get());
  S->active.clear();
  for (hipStream_t l : S->lanes) (void)hipStreamSynchronize(l);
  for (Direct& d : S->directs) {
    d.inst->free_buffers();
    S->free_events.push_back(d.ev);
  }
  S->directs.clear();
  for (hipStream_t l : S->lanes) (void)hipStreamDestroy(l);
  S->lanes.clear();
  for (hipEvent_t e : S->free_events) (void)hipEventDestroy(e);
  S->free_events.clear();
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
  uint32_t npts = 0, mlen = 0;
  int rc = validate_args(*args, t, &npts, &mlen);
  if (rc) return rc;
  if (args->refold_rounds)
    return fail(ZS_EINVAL, "served proofs fold every round (refold_rounds must be 0)");
  if (args->mode == ZS_MODE_LOGUP)
    return fail(ZS_EINVAL, "LogUp stages are proved one at a time (zs_prove)");
  const bool eq = args->mode == ZS_MODE_EQ;
  const uint32_t cb = field_bytes(args->chal_field);
  Pending p;
  p.a = *args;
  p.a.wire = nullptr;  // re-pointed at the copy on admission
  p.t = *t;
  p.mlen = mlen;
  if (args->wire) {
    p.has_wire = true;
    p.wire = *args->wire;
    if (p.wire.kind == ZS_WIRE_DOMAIN && p.wire.num_nodes)
      p.nodes.assign((const uint8_t*)p.wire.nodes,
                     (const uint8_t*)p.wire.nodes + (size_t)p.wire.num_nodes * cb);
  }
  p.claim.assign((const uint8_t*)args->claim, (const uint8_t*)args->claim + cb);
  if (eq) p.eq_point.assign((const uint8_t*)args->eq_point,
                            (const uint8_t*)args->eq_point + (size_t)args->rounds * cb);
  if (eq && args->eq_mass)
    p.eq_mass.assign((const uint8_t*)args->eq_mass, (const uint8_t*)args->eq_mass + cb);
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
  p.npts = npts;
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
