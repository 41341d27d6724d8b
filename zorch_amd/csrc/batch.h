// Batched proving: many independent proofs of one kind (field pair, factor
// count, mode) advance round by round together. Step s is one launch of
// batch_kernel that runs round s of every member still running, each member
// on a power-of-two share of the grid sized to its round; a member whose
// round has at most tail_pairs pairs gets one block that runs all its
// remaining rounds. The per-round latency chain (reduction, Fiat-Shamir hop,
// launch) is then paid once per step for the whole batch instead of once per
// round per proof, and the finishers of different members run on different
// CUs at the same time. A member whose round 2 is a pass round takes its
// first three rounds from the input, as a single proof does (lookahead on):
// step 0 runs its rounds 0 and 1 (batch01_kernel), step 1 its round 2
// (batch2_kernel), and from step 2 on it is one round ahead of the step.
//
// Device block layout (one pool allocation):
//   io records (one per member: control block | messages | raw point | final
//   values | eq point | eq line coefficients) | StageArgs array | step items |
//   eq family jobs            -- one host-to-device copy from pinned staging
//   slot rows (zeroed on the device) | per member: eq_l, eq_r families,
//   folded tables, staged input tables
// The io records come back with one device-to-host copy.
#pragma once
#include <algorithm>
#include <vector>

#include "instance.h"

namespace zs {

template <class FT, class FC, int K, int NPTS, bool EQ>
class BatchGroupImpl final : public BatchGroup {
  using TT = typename FT::T;
  using TC = typename FC::T;
  // Slot rows of the passes, then the look-ahead rows rounds 0-1 sum into.
  static constexpr size_t SLOT_BYTES = ((size_t)NSLOT * SLOTW + LA_SLOT_WORDS) * sizeof(uint64_t);

  struct Member {
    size_t index = 0;
    zs_prove_args a{};
    zs_transcript t{};
    uint32_t n = 0;
    size_t rec = 0, fam_l = 0, fam_r = 0, work = 0, in_stage = 0;
  };
  // One launch of a step: kind 0 = batch_kernel (passes of one round type and
  // tails), 1 = batch01_kernel (rounds 0-1 from the input), 2 = batch2_kernel
  // (round 2 from the input).
  struct Launch {
    uint32_t kind = 0;
    size_t item_off = 0;
    uint32_t nitems = 0;
    uint32_t blocks = 0;
  };
  struct Step {
    std::vector<Launch> launches;
    uint64_t bytes = 0;  // algorithmic bytes of the step (fused-pass model)
  };

  // Offsets inside one io record of a member with n rounds.
  static size_t o_msgs() { return align256(sizeof(Ctl<FC>)); }
  static size_t o_point(uint32_t n) { return o_msgs() + align256((size_t)n * NPTS * sizeof(TC)); }
  static size_t o_final(uint32_t n) { return o_point(n) + align256((size_t)n * sizeof(TC)); }
  static size_t o_eqp(uint32_t n) { return o_final(n) + align256((size_t)(K + 1) * sizeof(TC)); }
  static size_t o_eqab(uint32_t n) { return o_eqp(n) + (EQ ? align256((size_t)n * sizeof(TC)) : 0); }
  static size_t rec_bytes(uint32_t n) {
    return o_eqab(n) + (EQ ? align256((size_t)2 * n * sizeof(TC)) : 0);
  }

  // Algorithmic HBM bytes of round r of a member with n rounds.
  static uint64_t pass_bytes(uint32_t n, uint32_t r) {
    const uint64_t wt = sizeof(TT), wc = sizeof(TC);
    const uint64_t len = (uint64_t)1 << (n - r);
    if (r == 0) return (uint64_t)K * len * wt;
    const uint64_t w_prev = (r == 1) ? wt : wc;
    return (uint64_t)K * (2 * len) * w_prev + (uint64_t)K * len * wc;
  }

 public:
  ~BatchGroupImpl() override { free_buffers(); }

  void add(size_t index, const zs_prove_args& a, const zs_transcript& t) override {
    Member m;
    m.index = index;
    m.a = a;
    m.t = t;
    m.n = a.rounds;
    members_.push_back(m);
  }

  uint32_t max_rounds() const override { return max_n_; }

  int setup(zs_ctx* ctx, hipStream_t s) override {
    ctx_ = ctx;
    tune_ = ctx->tune;
    const size_t count = members_.size();
    size_t off = 0;
    for (auto& m : members_) {
      m.rec = off;
      off += rec_bytes(m.n);
      if (m.n > max_n_) max_n_ = m.n;
    }
    rec_total_ = off;
    args_off_ = off;
    off = align256(off + count * sizeof(StageArgs<FT, FC>));
    build_schedule();
    items_off_ = off;
    off = align256(off + items_.size() * sizeof(BatchItem));
    jobs_off_ = off;
    if (EQ) off = align256(off + count * sizeof(EqFamJob<FC>));
    h2d_bytes_ = off;
    slots_off_ = off;
    off = align256(off + count * SLOT_BYTES);
    for (auto& m : members_) {
      const size_t N = (size_t)1 << m.n;
      const uint32_t l_half = m.n / 2;
      if (EQ) {
        m.fam_l = off;
        off = align256(off + ((size_t)2 << l_half) * sizeof(TC));
        m.fam_r = off;
        off = align256(off + ((size_t)2 << (m.n - l_half)) * sizeof(TC));
      }
      m.work = off;
      off = align256(off + (size_t)K * (N / 2 > 0 ? N / 2 : 1) * sizeof(TC));
      if (!m.a.tables_on_device) {
        m.in_stage = off;
        off = align256(off + (size_t)K * N * sizeof(TT));
      }
    }
    void* block = nullptr;
    int rc = ctx->pool->alloc(off, &block);
    if (rc) return rc;
    block_ = block;
    device_bytes = off;
    if ((rc = ctx->host_pool->alloc(h2d_bytes_, &host_))) return rc;
    uint8_t* b = (uint8_t*)block;
    uint8_t* h = (uint8_t*)host_;
    memset(h, 0, h2d_bytes_);
    Ctl<FC> proto;
    memset(&proto, 0, sizeof(proto));
    init_ctl_consts<FC, K, NPTS>(proto);
    StageArgs<FT, FC>* args = (StageArgs<FT, FC>*)(h + args_off_);
    EqFamJob<FC>* jobs = (EqFamJob<FC>*)(h + jobs_off_);
    uint32_t eq_blocks = 0;
    for (size_t i = 0; i < count; ++i) {
      Member& m = members_[i];
      uint8_t* rec = h + m.rec;
      Ctl<FC>& c = *(Ctl<FC>*)rec;
      c = proto;
      if ((rc = set_ctl_proof<FC>(c, m.t, m.a, NPTS))) return rc;
      if (EQ) fill_eq_point<FC>((TC*)(rec + o_eqp(m.n)), (TC*)(rec + o_eqab(m.n)), m.a.eq_point, m.n);
      const size_t N = (size_t)1 << m.n;
      const uint32_t l_half = m.n / 2;
      StageArgs<FT, FC>& sa = args[i];
      sa.in = m.a.tables_on_device ? (const TT*)m.a.tables : (const TT*)(b + m.in_stage);
      sa.in_stride = N;
      sa.work = (TC*)(b + m.work);
      sa.work_stride = N / 2 > 0 ? N / 2 : 1;
      sa.n = m.n;
      sa.k = K;
      sa.ctl = (Ctl<FC>*)(b + m.rec);
      sa.slots = (uint64_t*)(b + slots_off_ + i * SLOT_BYTES);
      sa.msgs = (TC*)(b + m.rec + o_msgs());
      sa.point = b + m.rec + o_point(m.n);
      sa.final_out = (TC*)(b + m.rec + o_final(m.n));
      sa.eq_point = EQ ? (const TC*)(b + m.rec + o_eqp(m.n)) : nullptr;
      sa.eq_ab = EQ ? (const TC*)(b + m.rec + o_eqab(m.n)) : nullptr;
      sa.eq_l_fam = EQ ? (const TC*)(b + m.fam_l) : nullptr;
      sa.eq_r_fam = EQ ? (const TC*)(b + m.fam_r) : nullptr;
      sa.trace = ctx->trace;
      if (EQ) {
        EqFamJob<FC>& j = jobs[i];
        j.v = sa.eq_point;
        j.fam_l = (TC*)(b + m.fam_l);
        j.fam_r = (TC*)(b + m.fam_r);
        j.h_l = l_half;
        j.h_r = m.n - l_half;
        j.block_off = eq_blocks;
        const size_t entries = fam_offset(l_half) + fam_offset(m.n - l_half);
        eq_blocks += (uint32_t)((entries + 255) / 256);
      }
    }
    memcpy(h + items_off_, items_.data(), items_.size() * sizeof(BatchItem));
    for (auto& m : members_) {
      if (!m.a.tables_on_device) {
        const size_t N = (size_t)1 << m.n;
        ZS_HIP(hipMemcpyAsync(b + m.in_stage, m.a.tables, (size_t)K * N * sizeof(TT),
                              hipMemcpyHostToDevice, s));
      }
    }
    ZS_HIP(hipMemcpyAsync(b, h, h2d_bytes_, hipMemcpyHostToDevice, s));
    ZS_HIP(hipMemsetAsync(b + slots_off_, 0, count * SLOT_BYTES, s));
    if (EQ && eq_blocks) {
      eq_family_batch_kernel<FC><<<eq_blocks, 256, 0, s>>>((const EqFamJob<FC>*)(b + jobs_off_),
                                                           (uint32_t)count);
      ZS_HIP(hipGetLastError());
    }
    return ZS_OK;
  }

  int enqueue_step(uint32_t st, hipStream_t s) override {
    if (st >= steps_.size()) return ZS_OK;
    const Step& S = steps_[st];
    if (S.launches.empty()) return ZS_OK;
    uint8_t* b = (uint8_t*)block_;
    const StageArgs<FT, FC>* args = (const StageArgs<FT, FC>*)(b + args_off_);
    hipEvent_t ev = nullptr;
    if (ctx_->profile) {
      ev = ctx_->event();
      (void)hipEventRecord(ev, s);
    }
    uint32_t blocks = 0;
    for (const Launch& L : S.launches) {
      const BatchItem* items = (const BatchItem*)(b + items_off_) + L.item_off;
      blocks += L.blocks;
      if (L.kind == 1) {
        batch01_kernel<FT, FC, K, EQ><<<L.blocks, PASS_THREADS, 0, s>>>(args, items, L.nitems);
      } else if (L.kind == 2) {
        batch2_kernel<FT, FC, K, EQ><<<L.blocks, PASS_THREADS, 0, s>>>(args, items, L.nitems);
      } else if (st == 0) {
        batch_kernel<FT, FC, K, EQ, 0><<<L.blocks, PASS_THREADS, 0, s>>>(args, items, L.nitems);
      } else if (st == 1) {
        batch_kernel<FT, FC, K, EQ, 1><<<L.blocks, PASS_THREADS, 0, s>>>(args, items, L.nitems);
      } else {
        batch_kernel<FT, FC, K, EQ, 2><<<L.blocks, PASS_THREADS, 0, s>>>(args, items, L.nitems);
      }
      ZS_HIP(hipGetLastError());
    }
    if (ev) {
      hipEvent_t stop = ctx_->event();
      (void)hipEventRecord(stop, s);
      ctx_->records.push_back(KernelRecord{st, KREC_BATCH, blocks, S.bytes, ev, stop});
    }
    return ZS_OK;
  }

  int download(hipStream_t s) override {
    ZS_HIP(hipMemcpyAsync(host_, block_, rec_total_, hipMemcpyDeviceToHost, s));
    return ZS_OK;
  }

  int collect(zs_proof* outs, zs_transcript* ts) override {
    const uint8_t* h = (const uint8_t*)host_;
    for (const Member& m : members_) {
      const uint8_t* rec = h + m.rec;
      const Ctl<FC>& c = *(const Ctl<FC>*)rec;
      zs_proof& o = outs[m.index];
      if (o.msgs) copy_msg_rows(o.msgs, rec + o_msgs(), m.n, NPTS, c.mlen, sizeof(TC));
      if (o.point) memcpy(o.point, rec + o_point(m.n), (size_t)m.n * sizeof(TC));
      if (o.value) memcpy(o.value, &c.claim, sizeof(TC));
      if (o.final_tables)
        memcpy(o.final_tables, rec + o_final(m.n), (K + (EQ ? 1 : 0)) * sizeof(TC));
      if (ts) memcpy(&ts[m.index].sha, &c.tr, sizeof(TrState));
    }
    return ZS_OK;
  }

  void free_buffers() override {
    if (block_ && ctx_) ctx_->pool->release(block_);
    if (host_ && ctx_) ctx_->host_pool->release(host_);
    block_ = nullptr;
    host_ = nullptr;
  }

 private:
  // Items of every step: the next round of each running member, or the next
  // run of at most step_rounds tail rounds (all of them when 0); tails first
  // (they run longest), then passes by decreasing share of the grid. The
  // passes of a plain launch all have the step's round type (round 0, round 1,
  // or later), since a member taking its first rounds from the input runs
  // rounds 0-1 in step 0 and round 2 in step 1 in launches of their own.
  void build_schedule() {
    std::vector<uint32_t> next(members_.size(), 0);
    const uint32_t maxg = tune_.max_grid < (uint32_t)MAXG ? tune_.max_grid : (uint32_t)MAXG;
    const size_t per_block = (size_t)PASS_THREADS * tune_.pairs_per_thread;
    auto grid_for = [&](size_t work) {
      size_t grid = (work + per_block - 1) / per_block;
      if (grid > maxg) grid = maxg;
      if (grid < 1) grid = 1;
      while (grid & (grid - 1)) grid &= grid - 1;
      return (uint32_t)grid;
    };
    std::vector<bool> fused(members_.size(), false);
    for (size_t i = 0; i < members_.size(); ++i) {
      const uint32_t n = members_[i].n;
      fused[i] = tune_.lookahead && n >= 4 && ((size_t)1 << (n - 3)) > tune_.tail_pairs;
    }
    for (uint32_t st = 0; st < max_n_; ++st) {
      std::vector<BatchItem> tails, passes, first, second;
      Step S;
      for (size_t i = 0; i < members_.size(); ++i) {
        const uint32_t n = members_[i].n, round = next[i];
        if (round >= n) continue;
        const size_t N = (size_t)1 << n;
        if (fused[i] && round == 0) {
          // Rounds 0-1 stream 16-byte vectors of groups, as a single proof does.
          first.push_back(BatchItem{(uint32_t)i, grid_for(N / 4 / (16 / sizeof(TT))), 0, 0});
          next[i] = 2;
          S.bytes += (uint64_t)K * N * sizeof(TT);
          continue;
        }
        if (fused[i] && round == 2) {
          second.push_back(BatchItem{(uint32_t)i, grid_for(N / 8), 0, 2});
          next[i] = 3;
          S.bytes += (uint64_t)K * N * sizeof(TT) + (uint64_t)K * (N / 4) * sizeof(TC);
          continue;
        }
        const size_t pairs = (size_t)1 << (n - round - 1);
        if (pairs <= tune_.tail_pairs) {
          const uint32_t end =
              tune_.step_rounds && n - round > tune_.step_rounds ? round + tune_.step_rounds : n;
          tails.push_back(BatchItem{(uint32_t)i, 1, 0, tail_item(round, end)});
          next[i] = end;
          for (uint32_t r = round; r < end; ++r) S.bytes += pass_bytes(n, r);
          continue;
        }
        passes.push_back(BatchItem{(uint32_t)i, grid_for(pairs), 0, round});
        next[i] = round + 1;
        S.bytes += pass_bytes(n, round);
      }
      auto by_share = [](const BatchItem& x, const BatchItem& y) { return x.blocks > y.blocks; };
      std::stable_sort(passes.begin(), passes.end(), by_share);
      std::stable_sort(first.begin(), first.end(), by_share);
      std::stable_sort(second.begin(), second.end(), by_share);
      auto add_launch = [&](uint32_t kind, std::initializer_list<std::vector<BatchItem>*> lists) {
        Launch L;
        L.kind = kind;
        L.item_off = items_.size();
        uint32_t off = 0;
        for (auto* list : lists) {
          for (BatchItem it : *list) {
            it.block_off = off;
            off += it.blocks;
            items_.push_back(it);
          }
        }
        L.nitems = (uint32_t)(items_.size() - L.item_off);
        L.blocks = off;
        if (L.nitems) S.launches.push_back(L);
      };
      add_launch(0, {&tails, &passes});
      add_launch(1, {&first});
      add_launch(2, {&second});
      steps_.push_back(S);
    }
  }

  zs_ctx* ctx_ = nullptr;
  TuneParams tune_;
  std::vector<Member> members_;
  std::vector<BatchItem> items_;
  std::vector<Step> steps_;
  uint32_t max_n_ = 0;
  size_t rec_total_ = 0, args_off_ = 0, items_off_ = 0, jobs_off_ = 0, h2d_bytes_ = 0,
         slots_off_ = 0;
  void* block_ = nullptr;
  void* host_ = nullptr;
};

// The BatchGroupImpl of K factor tables for a message of npts values (K + 1, or K + 2 with the eq factor).
template <class FT, class FC, int K>
BatchGroup* make_typed_batch_k(uint32_t npts, bool eq) {
  if (eq && npts == (uint32_t)(K + 2)) return new BatchGroupImpl<FT, FC, K, K + 2, true>();
  if (!eq && npts == (uint32_t)(K + 1)) return new BatchGroupImpl<FT, FC, K, K + 1, false>();
  return nullptr;
}

}  // namespace zs
