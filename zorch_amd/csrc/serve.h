// Continuous-batching proof server, typed part: one ServeKindImpl per kind of
// request (field pair, factor count, mode). A request is admitted with its
// own device block (control block, outputs, slot rows, eq families, folded
// tables) and then advances one round per server step; the rounds that all
// admitted requests of a kind have pending are batched into one launch of
// batch_kernel with the round type chosen per item, so requests that arrive
// while others are mid-proof join at the next step. A request whose round
// has at most tail_pairs pairs runs all its remaining rounds in one block of
// that step and completes with it. A request whose round 2 is a pass round
// takes its first three rounds from the input (lookahead on): rounds 0-1 in
// one step (batch01_kernel), round 2 in the next (batch2_kernel).
//
// Request device block: io record (control block | messages | raw point |
// final values | eq point | eq line coefficients) | slot rows | eq_l, eq_r
// families | folded tables | staged input tables.
#pragma once
#include <algorithm>
#include <vector>

#include "instance.h"

namespace zs {

template <class FT, class FC, int K, int NPTS, bool EQ>
class ServeKindImpl final : public ServeKind {
  using TT = typename FT::T;
  using TC = typename FC::T;
  // Slot rows of the passes, then the look-ahead rows rounds 0-1 sum into.
  static constexpr size_t SLOT_BYTES = ((size_t)NSLOT * SLOTW + LA_SLOT_WORDS) * sizeof(uint64_t);

  struct Req final : ServeReq {
    bool fused = false;  // rounds 0-2 from the input
    void* block = nullptr;
    void* host = nullptr;  // pinned mirror of the io record
    size_t rec_bytes = 0, out_bytes = 0;
    StageArgs<FT, FC> sa{};
  };

  static size_t o_msgs() { return align256(sizeof(Ctl<FC>)); }
  static size_t o_point(uint32_t n) { return o_msgs() + align256((size_t)n * NPTS * sizeof(TC)); }
  static size_t o_final(uint32_t n) { return o_point(n) + align256((size_t)n * sizeof(TC)); }
  static size_t o_eqp(uint32_t n) { return o_final(n) + align256((size_t)(K + 1) * sizeof(TC)); }
  static size_t o_eqab(uint32_t n) { return o_eqp(n) + (EQ ? align256((size_t)n * sizeof(TC)) : 0); }
  static size_t rec_bytes(uint32_t n) {
    return o_eqab(n) + (EQ ? align256((size_t)2 * n * sizeof(TC)) : 0);
  }

  // Device block layout of a request with n rounds; returns its size.
  static size_t layout(uint32_t n, bool stage_in, size_t* o_slots, size_t* o_fl, size_t* o_fr,
                       size_t* o_work, size_t* o_in) {
    const size_t N = (size_t)1 << n;
    const uint32_t l_half = n / 2;
    size_t off = align256(rec_bytes(n));
    *o_slots = off;
    off = align256(off + SLOT_BYTES);
    *o_fl = off;
    if (EQ) off = align256(off + ((size_t)2 << l_half) * sizeof(TC));
    *o_fr = off;
    if (EQ) off = align256(off + ((size_t)2 << (n - l_half)) * sizeof(TC));
    *o_work = off;
    off = align256(off + (size_t)K * (N / 2 > 0 ? N / 2 : 1) * sizeof(TC));
    *o_in = off;
    if (stage_in) off = align256(off + (size_t)K * N * sizeof(TT));
    return off;
  }

 public:
  ~ServeKindImpl() override {
    for (Stage& st : ring_) {
      if (st.ev) {
        (void)hipEventSynchronize(st.ev);
        (void)hipEventDestroy(st.ev);
      }
      if (ctx_) {
        if (st.dev) ctx_->pool->release(st.dev);
        if (st.host) ctx_->host_pool->release(st.host);
      }
    }
  }

  uint64_t bytes_needed(const zs_prove_args& a) const override {
    size_t s, fl, fr, w, in;
    return layout(a.rounds, !a.tables_on_device, &s, &fl, &fr, &w, &in);
  }

  int admit(zs_ctx* ctx, const zs_prove_args& a, const zs_transcript& t, ServeReq** out) override {
    ctx_ = ctx;
    auto r = std::make_unique<Req>();
    const uint32_t n = a.rounds;
    const size_t N = (size_t)1 << n;
    const uint32_t l_half = n / 2;
    size_t o_slots, o_fl, o_fr, o_work, o_in;
    const bool stage_in = !a.tables_on_device;
    const size_t bytes = layout(n, stage_in, &o_slots, &o_fl, &o_fr, &o_work, &o_in);
    int rc = ctx->pool->alloc(bytes, &r->block);
    if (rc) return rc;
    r->rec_bytes = rec_bytes(n);
    r->out_bytes = o_eqp(n);
    if ((rc = ctx->host_pool->alloc(r->rec_bytes, &r->host))) {
      ctx->pool->release(r->block);
      return rc;
    }
    r->n = n;
    r->next = 0;
    r->bytes = bytes;
    r->fused = ctx->tune.lookahead && n >= 4 && ((size_t)1 << (n - 3)) > ctx->tune.tail_pairs;
    uint8_t* b = (uint8_t*)r->block;
    uint8_t* h = (uint8_t*)r->host;
    memset(h, 0, r->rec_bytes);
    Ctl<FC>& c = *(Ctl<FC>*)h;
    init_ctl_consts<FC, K, NPTS>(c);
    if ((rc = set_ctl_proof<FC>(c, t, a, NPTS))) {
      ctx->pool->release(r->block);
      ctx->host_pool->release(r->host);
      return rc;
    }
    if (EQ) fill_eq_point<FC>((TC*)(h + o_eqp(n)), (TC*)(h + o_eqab(n)), a.eq_point, n);
    StageArgs<FT, FC>& sa = r->sa;
    sa.in = stage_in ? (const TT*)(b + o_in) : (const TT*)a.tables;
    sa.in_stride = N;
    sa.work = (TC*)(b + o_work);
    sa.work_stride = N / 2 > 0 ? N / 2 : 1;
    sa.n = n;
    sa.k = K;
    sa.ctl = (Ctl<FC>*)b;
    sa.slots = (uint64_t*)(b + o_slots);
    sa.msgs = (TC*)(b + o_msgs());
    sa.point = b + o_point(n);
    sa.final_out = (TC*)(b + o_final(n));
    sa.eq_point = EQ ? (const TC*)(b + o_eqp(n)) : nullptr;
    sa.eq_ab = EQ ? (const TC*)(b + o_eqab(n)) : nullptr;
    sa.eq_l_fam = EQ ? (const TC*)(b + o_fl) : nullptr;
    sa.eq_r_fam = EQ ? (const TC*)(b + o_fr) : nullptr;
    sa.trace = ctx->trace;
    hipStream_t s = ctx->stream;
    if (stage_in)
      ZS_HIP(hipMemcpyAsync(b + o_in, a.tables, (size_t)K * N * sizeof(TT), hipMemcpyHostToDevice, s));
    ZS_HIP(hipMemcpyAsync(b, h, r->rec_bytes, hipMemcpyHostToDevice, s));
    ZS_HIP(hipMemsetAsync(b + o_slots, 0, SLOT_BYTES, s));
    if (EQ) {
      const size_t entries = fam_offset(l_half) + fam_offset(n - l_half);
      eq_family_kernel<FC><<<(entries + 255) / 256, 256, 0, s>>>(sa.eq_point, l_half,
                                                                   (TC*)sa.eq_l_fam, n - l_half,
                                                                   (TC*)sa.eq_r_fam);
      ZS_HIP(hipGetLastError());
    }
    *out = r.release();
    return ZS_OK;
  }

  int step(zs_ctx* ctx, std::vector<ServeReq*>& reqs, hipStream_t s) override {
    const size_t m = reqs.size();
    if (m == 0) return ZS_OK;
    const TuneParams& tune = ctx->tune;
    const uint32_t maxg = tune.max_grid < (uint32_t)MAXG ? tune.max_grid : (uint32_t)MAXG;
    // Pass items by round type (round 0, round 1, later rounds), one launch
    // each: a launch's kernel is specialised to its round type. Tail items
    // ride with the later rounds.
    std::vector<BatchItem> tails, passes[3], first01, second2;
    const size_t per_block = (size_t)PASS_THREADS * tune.pairs_per_thread;
    auto grid_for = [&](size_t work) {
      size_t grid = (work + per_block - 1) / per_block;
      if (grid > maxg) grid = maxg;
      if (grid < 1) grid = 1;
      while (grid & (grid - 1)) grid &= grid - 1;
      return (uint32_t)grid;
    };
    for (size_t i = 0; i < m; ++i) {
      Req* r = static_cast<Req*>(reqs[i]);
      const uint32_t round = r->next;
      const size_t pairs = (size_t)1 << (r->n - round - 1);
      const size_t N = (size_t)1 << r->n;
      if (r->fused && round == 0) {
        first01.push_back(BatchItem{(uint32_t)i, grid_for(N / 4 / (16 / sizeof(TT))), 0, 0});
        r->next = 2;
      } else if (r->fused && round == 2) {
        second2.push_back(BatchItem{(uint32_t)i, grid_for(N / 8), 0, 2});
        r->next = 3;
      } else if (pairs <= tune.tail_pairs) {
        // A run of at most step_rounds tail rounds (all of them when 0).
        const uint32_t end = tune.step_rounds && r->n - round > tune.step_rounds
                                 ? round + tune.step_rounds
                                 : r->n;
        tails.push_back(BatchItem{(uint32_t)i, 1, 0, tail_item(round, end)});
        r->next = end;
      } else {
        passes[round < 2 ? round : 2].push_back(BatchItem{(uint32_t)i, grid_for(pairs), 0, round});
        r->next = round + 1;
      }
      r->finishing = r->next == r->n;
    }
    for (auto* list : {&passes[0], &passes[1], &passes[2], &first01, &second2})
      std::stable_sort(list->begin(), list->end(),
                       [](const BatchItem& x, const BatchItem& y) { return x.blocks > y.blocks; });
    const size_t args_bytes = align256(m * sizeof(StageArgs<FT, FC>));
    const size_t need = args_bytes + m * sizeof(BatchItem);  // one item a request
    // The next staging slot of the ring, free once the step that last used
    // it has run (steps are queued ahead of the device).
    Stage& st = ring_[ring_next_];
    ring_next_ = (ring_next_ + 1) % NSTAGE;
    if (st.ev) ZS_HIP(hipEventSynchronize(st.ev));
    else ZS_HIP(hipEventCreateWithFlags(&st.ev, hipEventDisableTiming));
    if (need > st.cap) {
      if (st.dev) ctx->pool->release(st.dev);
      if (st.host) ctx->host_pool->release(st.host);
      st.dev = st.host = nullptr;
      st.cap = 0;
      size_t cap = 4096;
      while (cap < need) cap *= 2;
      int rc = ctx->pool->alloc(cap, &st.dev);
      if (rc) return rc;
      if ((rc = ctx->host_pool->alloc(cap, &st.host))) return rc;
      st.cap = cap;
    }
    uint8_t* h = (uint8_t*)st.host;
    StageArgs<FT, FC>* args = (StageArgs<FT, FC>*)h;
    for (size_t i = 0; i < m; ++i) args[i] = static_cast<Req*>(reqs[i])->sa;
    BatchItem* items = (BatchItem*)(h + args_bytes);
    // Launch t takes items [first[t], first[t] + count[t]) over blocks[t] blocks:
    // round 0, round 1, later rounds and tails, rounds 0-1 from the input,
    // round 2 from the input.
    uint32_t first[5] = {}, count[5] = {}, blocks[5] = {};
    uint32_t nitems = 0;
    for (int t = 0; t < 5; ++t) {
      first[t] = nitems;
      uint32_t off = 0;
      std::vector<BatchItem>* lists[2] = {t == 3 ? &first01 : t == 4 ? &second2 : &passes[t],
                                          t == 2 ? &tails : nullptr};
      for (auto* list : lists) {
        if (!list) continue;
        for (BatchItem it : *list) {
          it.block_off = off;
          off += it.blocks;
          items[nitems++] = it;
        }
      }
      count[t] = nitems - first[t];
      blocks[t] = off;
    }
    ZS_HIP(hipMemcpyAsync(st.dev, h, need, hipMemcpyHostToDevice, s));
    uint8_t* d = (uint8_t*)st.dev;
    const StageArgs<FT, FC>* dargs = (const StageArgs<FT, FC>*)d;
    const BatchItem* ditems = (const BatchItem*)(d + args_bytes);
    if (count[0])
      batch_kernel<FT, FC, K, EQ, 0><<<blocks[0], PASS_THREADS, 0, s>>>(dargs, ditems + first[0],
                                                                        count[0]);
    if (count[1])
      batch_kernel<FT, FC, K, EQ, 1><<<blocks[1], PASS_THREADS, 0, s>>>(dargs, ditems + first[1],
                                                                        count[1]);
    if (count[2])
      batch_kernel<FT, FC, K, EQ, 2><<<blocks[2], PASS_THREADS, 0, s>>>(dargs, ditems + first[2],
                                                                        count[2]);
    if (count[3])
      batch01_kernel<FT, FC, K, EQ><<<blocks[3], PASS_THREADS, 0, s>>>(dargs, ditems + first[3],
                                                                       count[3]);
    if (count[4])
      batch2_kernel<FT, FC, K, EQ><<<blocks[4], PASS_THREADS, 0, s>>>(dargs, ditems + first[4],
                                                                      count[4]);
    ZS_HIP(hipGetLastError());
    ZS_HIP(hipEventRecord(st.ev, s));
    return ZS_OK;
  }

  int download(ServeReq* req, hipStream_t s) override {
    Req* r = static_cast<Req*>(req);
    ZS_HIP(hipMemcpyAsync(r->host, r->block, r->out_bytes, hipMemcpyDeviceToHost, s));
    return ZS_OK;
  }

  void collect(ServeReq* req, ServeResult* out) override {
    Req* r = static_cast<Req*>(req);
    const uint8_t* h = (const uint8_t*)r->host;
    const Ctl<FC>& c = *(const Ctl<FC>*)h;
    const uint32_t n = r->n;
    out->msgs.resize((size_t)n * c.mlen * sizeof(TC));
    copy_msg_rows(out->msgs.data(), h + o_msgs(), n, NPTS, c.mlen, sizeof(TC));
    out->point.assign(h + o_point(n), h + o_point(n) + (size_t)n * sizeof(TC));
    out->value.assign((const uint8_t*)&c.claim, (const uint8_t*)&c.claim + sizeof(TC));
    out->final_tables.assign(h + o_final(n), h + o_final(n) + (K + (EQ ? 1 : 0)) * sizeof(TC));
    out->tr_state.assign((const uint8_t*)&c.tr, (const uint8_t*)&c.tr + sizeof(TrState));
  }

  void release(zs_ctx* ctx, ServeReq* req) override {
    Req* r = static_cast<Req*>(req);
    if (r->block) ctx->pool->release(r->block);
    if (r->host) ctx->host_pool->release(r->host);
    r->block = r->host = nullptr;
  }

 private:
  // Staging of a step's argument records and items: one slot per step that
  // can be queued at once, plus one being filled.
  struct Stage {
    void* dev = nullptr;
    void* host = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
  };
  static constexpr int NSTAGE = 4;
  zs_ctx* ctx_ = nullptr;
  Stage ring_[NSTAGE];
  int ring_next_ = 0;
};

// The ServeKindImpl of K factor tables for a message of npts values (K + 1, or K + 2 with the eq factor).
template <class FT, class FC, int K>
ServeKind* make_typed_serve_k(uint32_t npts, bool eq) {
  if (eq && npts == (uint32_t)(K + 2)) return new ServeKindImpl<FT, FC, K, K + 2, true>();
  if (!eq && npts == (uint32_t)(K + 1)) return new ServeKindImpl<FT, FC, K, K + 1, false>();
  return nullptr;
}

}  // namespace zs
