// Typed proof instance: device buffers, uploads, the per-round launch
// sequence and the downloads. Instantiated per field combination in
// inst_*.hip (one translation unit each, compiled in parallel).
#pragma once
#include <string.h>

#include <vector>

#include "kernels.h"
#include "runtime.h"

namespace zs {

template <class F>
typename F::T load_host_el(const void* src, size_t i) {
  typename F::T v;
  memcpy(&v, (const uint8_t*)src + i * sizeof(typename F::T), sizeof(v));
  return F::canon(v);
}

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

template <class FT, class FC, int K, int NPTS, bool EQ>
class InstanceImpl final : public Instance {
  using TT = typename FT::T;
  using TC = typename FC::T;

 public:
  ~InstanceImpl() override { free_buffers(); }

  int setup(zs_ctx* ctx, const zs_prove_args& a, const zs_transcript& t) override {
    ctx_ = ctx;
    n_ = a.rounds;
    rounds = n_;
    npts = NPTS;
    k = K;
    chal_bytes = sizeof(TC);
    tune_ = ctx->tune;
    const size_t N = (size_t)1 << n_;
    const uint32_t l_half = n_ / 2;
    // Carve one block.
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off = align256(off + bytes); return o; };
    const size_t o_ctl = take(sizeof(Ctl<FC>));
    const size_t o_msgs = take((size_t)n_ * NPTS * sizeof(TC));
    const size_t o_point = take((size_t)n_ * sizeof(TC));
    const size_t o_final = take((size_t)(K + 1) * sizeof(TC));
    const size_t o_part = take((size_t)MAXG * NPTS * sizeof(TC));
    size_t o_eqp = 0, o_eql = 0, o_eqr = 0;
    if (EQ) {
      o_eqp = take((size_t)n_ * sizeof(TC));
      o_eql = take(((size_t)2 << l_half) * sizeof(TC));
      o_eqr = take(((size_t)2 << (n_ - l_half)) * sizeof(TC));
    }
    const size_t work_stride = N / 2 > 0 ? N / 2 : 1;
    const size_t o_work = take((size_t)K * work_stride * sizeof(TC));
    size_t o_in = 0;
    const bool stage_in = !a.tables_on_device;
    if (stage_in) o_in = take((size_t)K * N * sizeof(TT));
    void* block = nullptr;
    int rc = ctx->pool->alloc(off, &block);
    if (rc) return rc;
    bufs_.block = block;
    bufs_.block_bytes = off;
    device_bytes = off;
    uint8_t* b = (uint8_t*)block;
    args_.in = stage_in ? (const TT*)(b + o_in) : (const TT*)a.tables;
    args_.in_stride = N;
    args_.work = (TC*)(b + o_work);
    args_.work_stride = work_stride;
    args_.n = n_;
    args_.k = K;
    args_.ctl = (Ctl<FC>*)(b + o_ctl);
    args_.partials = (TC*)(b + o_part);
    args_.msgs = (TC*)(b + o_msgs);
    args_.point = b + o_point;
    args_.final_out = (TC*)(b + o_final);
    args_.eq_point = EQ ? (const TC*)(b + o_eqp) : nullptr;
    args_.eq_l_fam = EQ ? (const TC*)(b + o_eql) : nullptr;
    args_.eq_r_fam = EQ ? (const TC*)(b + o_eqr) : nullptr;
    hipStream_t s = ctx->stream;
    if (stage_in) {
      ZS_HIP(hipMemcpyAsync((void*)args_.in, a.tables, (size_t)K * N * sizeof(TT),
                            hipMemcpyHostToDevice, s));
    }
    // Control block.
    memset(&ctl_h_, 0, sizeof(ctl_h_));
    memcpy(&ctl_h_.sha, &t.sha, sizeof(Sha256State));
    ctl_h_.item_bytes = t.item_bytes;
    ctl_h_.limbs = (uint32_t)(sizeof(TC) / t.item_bytes);
    ctl_h_.claim = load_host_el<FC>(a.claim, 0);
    ctl_h_.r = FC::zero();
    ctl_h_.eq_prev = FC::one();
    ctl_h_.one = FC::one();
    for (int tt = 0; tt < NPTS; ++tt) {
      int64_t den = 1;
      for (int j = 0; j < NPTS; ++j)
        if (j != tt) den *= (int64_t)(tt - j);
      ctl_h_.lag_inv[tt] = FC::inv(FC::from_int(den));
    }
    ZS_HIP(hipMemcpyAsync(args_.ctl, &ctl_h_, sizeof(ctl_h_), hipMemcpyHostToDevice, s));
    if (EQ) {
      eq_h_.resize(n_);
      for (uint32_t i = 0; i < n_; ++i) eq_h_[i] = load_host_el<FC>(a.eq_point, i);
      ZS_HIP(hipMemcpyAsync((void*)args_.eq_point, eq_h_.data(), n_ * sizeof(TC),
                            hipMemcpyHostToDevice, s));
      if (l_half > 0)
        eq_family_kernel<FC><<<1, 256, 0, s>>>(args_.eq_point, l_half, (TC*)args_.eq_l_fam);
      if (n_ - l_half > 0)
        eq_family_kernel<FC><<<1, 256, 0, s>>>(args_.eq_point + l_half, n_ - l_half,
                                               (TC*)args_.eq_r_fam);
      ZS_HIP(hipGetLastError());
    }
    // Host staging for outputs.
    msgs_h_.resize((size_t)n_ * NPTS);
    point_h_.resize((size_t)n_ * sizeof(TC));
    final_h_.resize(K + 1);
    return ZS_OK;
  }

  // Algorithmic HBM bytes of round r's pass under the fused-pass model.
  uint64_t pass_bytes(uint32_t r) const {
    const uint64_t wt = sizeof(TT), wc = sizeof(TC);
    const uint64_t len = (uint64_t)1 << (n_ - r);
    if (r == 0) return (uint64_t)K * len * wt;
    const uint64_t w_prev = (r == 1) ? wt : wc;
    return (uint64_t)K * (2 * len) * w_prev + (uint64_t)K * len * wc;
  }

  void mark(hipStream_t s, uint32_t r, uint32_t kind, uint32_t grid, uint64_t bytes,
            hipEvent_t start) {
    hipEvent_t stop = ctx_->event();
    (void)hipEventRecord(stop, s);
    ctx_->records.push_back(KernelRecord{r, kind, grid, bytes, start, stop});
  }

  hipEvent_t begin(hipStream_t s) {
    if (!ctx_->profile) return nullptr;
    hipEvent_t e = ctx_->event();
    (void)hipEventRecord(e, s);
    return e;
  }

  int enqueue(uint32_t first, uint32_t last, hipStream_t s) override {
    for (uint32_t r = first; r < last; ++r) {
      const size_t pairs = (size_t)1 << (n_ - r - 1);
      if (pairs <= tune_.tail_pairs) {
        hipEvent_t ev = begin(s);
        tail_kernel<FT, FC, K, EQ><<<1, TAIL_THREADS, 0, s>>>(args_, r, last);
        ZS_HIP(hipGetLastError());
        if (ev) {
          uint64_t b = 0;
          for (uint32_t q = r; q < last; ++q) b += pass_bytes(q);
          mark(s, r, KREC_TAIL, 1, b, ev);
        }
        return ZS_OK;
      }
      size_t per_block = (size_t)PASS_THREADS * tune_.pairs_per_thread;
      size_t grid = (pairs + per_block - 1) / per_block;
      uint32_t maxg = tune_.max_grid < (uint32_t)MAXG ? tune_.max_grid : (uint32_t)MAXG;
      if (grid > maxg) grid = maxg;
      if (grid < 1) grid = 1;
      // A power-of-two grid keeps each thread's eq_r index fixed across its pairs.
      while (grid & (grid - 1)) grid &= grid - 1;
      const uint32_t L = 1u << (n_ - r);
      hipEvent_t ev = begin(s);
      if (r == 0) {
        pass_kernel<FT, FT, FC, K, EQ, false>
            <<<grid, PASS_THREADS, 0, s>>>(args_, args_.in, args_.in_stride, L, r);
      } else if (r == 1) {
        pass_kernel<FT, FT, FC, K, EQ, true>
            <<<grid, PASS_THREADS, 0, s>>>(args_, args_.in, args_.in_stride, L, r);
      } else {
        pass_kernel<FT, FC, FC, K, EQ, true>
            <<<grid, PASS_THREADS, 0, s>>>(args_, args_.work, args_.work_stride, L, r);
      }
      ZS_HIP(hipGetLastError());
      if (ev) mark(s, r, KREC_PASS, (uint32_t)grid, pass_bytes(r), ev);
    }
    return ZS_OK;
  }

  int download(hipStream_t s) override {
    ZS_HIP(hipMemcpyAsync(msgs_h_.data(), args_.msgs, msgs_h_.size() * sizeof(TC),
                          hipMemcpyDeviceToHost, s));
    ZS_HIP(hipMemcpyAsync(point_h_.data(), args_.point, point_h_.size(),
                          hipMemcpyDeviceToHost, s));
    ZS_HIP(hipMemcpyAsync(final_h_.data(), args_.final_out, final_h_.size() * sizeof(TC),
                          hipMemcpyDeviceToHost, s));
    ZS_HIP(hipMemcpyAsync(&ctl_h_, args_.ctl, sizeof(ctl_h_), hipMemcpyDeviceToHost, s));
    return ZS_OK;
  }

  int collect(zs_proof* out, zs_transcript* t) override {
    if (out->msgs) memcpy(out->msgs, msgs_h_.data(), msgs_h_.size() * sizeof(TC));
    if (out->point) memcpy(out->point, point_h_.data(), point_h_.size());
    if (out->value) memcpy(out->value, &ctl_h_.claim, sizeof(TC));
    if (out->final_tables) memcpy(out->final_tables, final_h_.data(), (K + (EQ ? 1 : 0)) * sizeof(TC));
    if (t) memcpy(&t->sha, &ctl_h_.sha, sizeof(Sha256State));
    return ZS_OK;
  }

  int upload_transcript(const zs_transcript& t, hipStream_t s) override {
    // The transcript prefix of the control block: state, word size, limbs.
    memcpy(&ctl_h_.sha, &t.sha, sizeof(Sha256State));
    ctl_h_.item_bytes = t.item_bytes;
    ctl_h_.limbs = (uint32_t)(sizeof(TC) / t.item_bytes);
    static_assert(offsetof(Ctl<FC>, limbs) + sizeof(uint32_t) == offsetof(Ctl<FC>, counter),
                  "control block prefix");
    ZS_HIP(hipMemcpyAsync(args_.ctl, &ctl_h_, offsetof(Ctl<FC>, counter), hipMemcpyHostToDevice, s));
    return ZS_OK;
  }

  int download_round(uint32_t round, hipStream_t s) override {
    ZS_HIP(hipMemcpyAsync(msgs_h_.data() + (size_t)round * NPTS, args_.msgs + (size_t)round * NPTS,
                          NPTS * sizeof(TC), hipMemcpyDeviceToHost, s));
    ZS_HIP(hipMemcpyAsync(point_h_.data() + (size_t)round * sizeof(TC),
                          args_.point + (size_t)round * sizeof(TC), sizeof(TC),
                          hipMemcpyDeviceToHost, s));
    ZS_HIP(hipMemcpyAsync(&ctl_h_, args_.ctl, sizeof(ctl_h_), hipMemcpyDeviceToHost, s));
    return ZS_OK;
  }

  int collect_round(uint32_t round, void* msg, void* chal, void* claim, zs_transcript* t) override {
    if (msg) memcpy(msg, msgs_h_.data() + (size_t)round * NPTS, NPTS * sizeof(TC));
    if (chal) memcpy(chal, point_h_.data() + (size_t)round * sizeof(TC), sizeof(TC));
    if (claim) memcpy(claim, &ctl_h_.claim, sizeof(TC));
    if (t) memcpy(&t->sha, &ctl_h_.sha, sizeof(Sha256State));
    return ZS_OK;
  }

  // Current tables after `bound` rounds, lifted into FC.
  int read_tables(uint32_t bound, void* out, size_t out_bytes, hipStream_t s) override {
    const size_t len = (size_t)1 << (n_ - bound);
    if (out_bytes < (size_t)K * len * sizeof(TC)) return fail(ZS_EINVAL, "output buffer too small");
    if (bound == n_) {
      ZS_HIP(hipMemcpyAsync(out, args_.final_out, (size_t)K * sizeof(TC), hipMemcpyDeviceToHost, s));
      ZS_HIP(hipStreamSynchronize(s));
      return ZS_OK;
    }
    if (bound == 0) {
      std::vector<TT> tmp((size_t)K * len);
      ZS_HIP(hipMemcpyAsync(tmp.data(), args_.in, tmp.size() * sizeof(TT), hipMemcpyDeviceToHost, s));
      ZS_HIP(hipStreamSynchronize(s));
      TC* o = (TC*)out;
      for (size_t i = 0; i < tmp.size(); ++i) o[i] = Lift<FT, FC>::go(tmp[i]);
      return ZS_OK;
    }
    for (uint32_t q = 0; q < K; ++q)
      ZS_HIP(hipMemcpyAsync((TC*)out + (size_t)q * len, args_.work + (size_t)q * args_.work_stride,
                            len * sizeof(TC), hipMemcpyDeviceToHost, s));
    ZS_HIP(hipStreamSynchronize(s));
    return ZS_OK;
  }

  void free_buffers() override {
    if (bufs_.block && ctx_) ctx_->pool->release(bufs_.block);
    bufs_.block = nullptr;
  }

 private:
  zs_ctx* ctx_ = nullptr;
  uint32_t n_ = 0;
  TuneParams tune_;
  InstanceBuffers bufs_;
  StageArgs<FT, FC> args_{};
  Ctl<FC> ctl_h_;
  Sha256State sha_h_;
  std::vector<TC> eq_h_;
  std::vector<TC> msgs_h_;
  std::vector<uint8_t> point_h_;
  std::vector<TC> final_h_;
};

template <class FT, class FC>
Instance* make_typed(uint32_t k, uint32_t npts, bool eq) {
#define ZS_CASE(KK, EQQ)                                                    \
  if (k == KK && eq == EQQ && npts == (uint32_t)(KK + 1 + (EQQ ? 1 : 0))) \
    return new InstanceImpl<FT, FC, KK, KK + 1 + (EQQ ? 1 : 0), EQQ>();
  ZS_CASE(1, false) ZS_CASE(2, false) ZS_CASE(3, false) ZS_CASE(4, false)
  ZS_CASE(1, true) ZS_CASE(2, true) ZS_CASE(3, true) ZS_CASE(4, true)
#undef ZS_CASE
  return nullptr;
}

}  // namespace zs
