// Typed proof instance: device buffers, uploads, the per-round launch
// sequence and the downloads. Instantiated per field combination in
// inst_<fields>_k<K>.hip (one translation unit per field pair and factor
// count, compiled in parallel).
//
// Device block layout (one pool allocation, 256-byte aligned regions):
//   io region:  control block | messages | raw point | final values | slot rows |
//               eq point | eq line coefficients
//   eq_l family | eq_r family | folded tables | staged input tables
// The io region is mirrored by one pinned host block, so a proof costs one
// host-to-device copy of the control block and eq point before its rounds and
// one device-to-host copy of the outputs after them.
#pragma once
#include <stddef.h>
#include <string.h>

#include <vector>

#include "engines.h"
#include "kernels.h"
#include "runtime.h"
#include "wire.h"

namespace zs {

template <class F>
typename F::T load_host_el(const void* src, size_t i) {
  typename F::T v;
  memcpy(&v, (const uint8_t*)src + i * sizeof(typename F::T), sizeof(v));
  return F::canon(v);
}

static_assert(TUNE_MAX_GRID == (uint32_t)MAXG, "grid limit");

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Interpolation constants of a control block over the nodes {0..NPTS-1}.
template <class FC, int K, int NPTS>
void init_ctl_consts(Ctl<FC>& c) {
  c.one = FC::one();
  for (int tt = 0; tt < NPTS; ++tt) {
    int64_t den = 1;
    for (int j = 0; j < NPTS; ++j)
      if (j != tt) den *= (int64_t)(tt - j);
    c.lag_inv[tt] = FC::inv(FC::from_int(den));
    c.nodes[tt] = FC::from_int(tt);
  }
  int64_t binom = 1;  // C(K+1, i)
  for (int i = 0; i <= K; ++i) {
    c.ext_coef[i] = FC::from_int(((K - i) & 1) ? -binom : binom);
    binom = binom * (K + 1 - i) / (i + 1);
  }
}

// The per-proof part of a control block: transcript, claim, running state,
// and the maps of the request's message wire form (ZS_OK or ZS_EINVAL).
template <class FC>
int set_ctl_proof(Ctl<FC>& c, const zs_transcript& t, const zs_prove_args& a, uint32_t npts) {
  static_assert(sizeof(TrState) == sizeof(zs_transcript) - offsetof(zs_transcript, sha),
                "transcript state layout");
  memcpy(&c.tr, &t.sha, sizeof(TrState));
  c.tr_kind = t.kind;
  c.item_bytes = t.item_bytes;
  c.limbs = (uint32_t)(sizeof(typename FC::T) / t.item_bytes);
  c.counter = 0;
  c.claim = load_host_el<FC>(a.claim, 0);
  c.r = FC::zero();
  c.eq_prev = a.eq_mass ? load_host_el<FC>(a.eq_mass, 0) : FC::one();
  return set_ctl_wire<FC>(c, npts, a.wire);
}

// Copy `rounds` message rows of the device layout (npts elements a row) into
// the caller's layout of mlen elements a row.
inline void copy_msg_rows(void* dst, const void* src, uint32_t rounds, uint32_t npts,
                          uint32_t mlen, size_t el_bytes) {
  for (uint32_t i = 0; i < rounds; ++i)
    memcpy((uint8_t*)dst + (size_t)i * mlen * el_bytes,
           (const uint8_t*)src + (size_t)i * npts * el_bytes, (size_t)mlen * el_bytes);
}

// Eq point and the coefficients of l_i(t) = eq_prev ((1 - w_i) + t (2 w_i - 1)).
template <class FC>
void fill_eq_point(typename FC::T* ep, typename FC::T* ab, const void* eq_point, uint32_t n) {
  for (uint32_t i = 0; i < n; ++i) {
    ep[i] = load_host_el<FC>(eq_point, i);
    ab[2 * i] = FC::sub(FC::add(ep[i], ep[i]), FC::one());  // 2 w_i - 1
    ab[2 * i + 1] = FC::sub(FC::one(), ep[i]);              // 1 - w_i
  }
}

template <class FT, class FC, int K, int NPTS, bool EQ>
class InstanceImpl final : public Instance {
  using TT = typename FT::T;
  using TC = typename FC::T;

 public:
  ~InstanceImpl() override { free_buffers(); }

  using Instance::setup;
  int setup(zs_ctx* ctx, const zs_prove_args& a, const zs_transcript& t, hipStream_t s,
            const TuneParams& tune) override {
    return setup_in(ctx, a, t, s, tune, nullptr, 0, nullptr, 0);
  }

  int setup_in(zs_ctx* ctx, const zs_prove_args& a, const zs_transcript& t, hipStream_t s,
               const TuneParams& tune, void* dev_mem, size_t dev_cap, void* host_mem,
               size_t host_cap) override {
    ctx_ = ctx;
    n_ = a.rounds;
    rounds = n_;
    npts = NPTS;
    k = K;
    chal_bytes = sizeof(TC);
    tune_ = tune;
    const size_t N = (size_t)1 << n_;
    const uint32_t l_half = n_ / 2;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off = align256(off + bytes); return o; };
    o_ctl_ = take(sizeof(Ctl<FC>));
    o_msgs_ = take((size_t)n_ * NPTS * sizeof(TC));
    o_point_ = take((size_t)n_ * sizeof(TC));
    o_final_ = take((size_t)(K + 1) * sizeof(TC));
    out_bytes_ = off;
    // Slot rows of the pass and chained kernels, then the look-ahead chain's.
    const size_t o_slots = take(((size_t)NSLOT * SLOTW + LA_SLOT_WORDS) * sizeof(uint64_t));
    size_t o_eqp = 0, o_eqab = 0, o_eql = 0, o_eqr = 0;
    if (EQ) {
      o_eqp = take((size_t)n_ * sizeof(TC));
      o_eqab = take((size_t)2 * n_ * sizeof(TC));
    }
    io_bytes_ = off;
    if (EQ) {
      o_eql = take(((size_t)2 << l_half) * sizeof(TC));
      o_eqr = take(((size_t)2 << (n_ - l_half)) * sizeof(TC));
    }
    const size_t work_stride = N / 2 > 0 ? N / 2 : 1;
    const size_t o_work = take((size_t)K * work_stride * sizeof(TC));
    // The sqrt-space engine's two eq(r[<i], .) buffers (2^h entries each).
    refold_ = a.refold_rounds;
    const size_t o_sq = refold_ ? take(((size_t)2 << refold_) * sizeof(TC)) : 0;
    size_t o_in = 0;
    const bool stage_in = !a.tables_on_device;
    if (stage_in) o_in = take((size_t)K * N * sizeof(TT));
    void* block = nullptr;
    int rc = ZS_OK;
    if (dev_mem) {
      // Caller-provided memory: nothing is allocated or released here.
      if (off > dev_cap)
        return fail(ZS_EINVAL, "the stage needs " + std::to_string(off) +
                                   " bytes of device memory; the buffer holds " +
                                   std::to_string(dev_cap));
      if (io_bytes_ > host_cap) return fail(ZS_EINVAL, "host staging buffer too small");
      external_ = true;
      block = dev_mem;
      host_ = host_mem;
    } else {
      if ((rc = ctx->pool->alloc(off, &block))) return rc;
    }
    block_ = block;
    device_bytes = off;
    if (!external_ && (rc = ctx->host_pool->alloc(io_bytes_, &host_))) return rc;
    uint8_t* b = (uint8_t*)block;
    args_.in = stage_in ? (const TT*)(b + o_in) : (const TT*)a.tables;
    args_.in_stride = N;
    args_.work = (TC*)(b + o_work);
    args_.work_stride = work_stride;
    args_.n = n_;
    args_.k = K;
    args_.ctl = (Ctl<FC>*)(b + o_ctl_);
    args_.slots = (uint64_t*)(b + o_slots);
    args_.msgs = (TC*)(b + o_msgs_);
    args_.point = b + o_point_;
    args_.final_out = (TC*)(b + o_final_);
    args_.eq_point = EQ ? (const TC*)(b + o_eqp) : nullptr;
    args_.eq_ab = EQ ? (const TC*)(b + o_eqab) : nullptr;
    args_.eq_l_fam = EQ ? (const TC*)(b + o_eql) : nullptr;
    args_.eq_r_fam = EQ ? (const TC*)(b + o_eqr) : nullptr;
    args_.trace = ctx->trace;
    args_.sq_eq = refold_ ? (TC*)(b + o_sq) : nullptr;
    args_.sq_half = refold_;
    if (stage_in) {
      ZS_HIP(hipMemcpyAsync((void*)args_.in, a.tables, (size_t)K * N * sizeof(TT),
                            hipMemcpyHostToDevice, s));
    }
    // Control block and eq point, staged in pinned memory, one copy.
    uint8_t* h = (uint8_t*)host_;
    memset(h, 0, io_bytes_);
    Ctl<FC>& c = *(Ctl<FC>*)(h + o_ctl_);
    init_ctl_consts<FC, K, NPTS>(c);
    if ((rc = set_ctl_proof<FC>(c, t, a, NPTS))) return rc;
    mlen = c.mlen;
    if (EQ) fill_eq_point<FC>((TC*)(h + o_eqp), (TC*)(h + o_eqab), a.eq_point, n_);
    ZS_HIP(hipMemcpyAsync(b, h, io_bytes_, hipMemcpyHostToDevice, s));
    if (EQ) {
      const size_t entries = fam_offset(l_half) + fam_offset(n_ - l_half);
      const size_t grid = (entries + 255) / 256;
      eq_family_kernel<FC><<<grid, 256, 0, s>>>(args_.eq_point, l_half, (TC*)args_.eq_l_fam,
                                                n_ - l_half, (TC*)args_.eq_r_fam);
      ZS_HIP(hipGetLastError());
    }
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

  // Whether round r of [.., last) starts a look-ahead chain in enqueue.
  bool chain_la_starts(uint32_t r, uint32_t last) const {
    if (!tune_.lookahead || r < 2 || r >= last) return false;
    const size_t pairs = (size_t)1 << (n_ - r - 1);
    if (pairs <= tune_.tail_pairs || pairs > tune_.chain_pairs) return false;
    uint32_t end = r;
    while (end < last && ((size_t)1 << (n_ - end - 1)) > tune_.tail_pairs) ++end;
    return end - r >= 2;
  }

  // Power-of-two grid for `units` thread units of work, as for a pass.
  size_t pass_grid(size_t units) const {
    const size_t per_block = (size_t)PASS_THREADS * tune_.pairs_per_thread;
    size_t grid = (units + per_block - 1) / per_block;
    const uint32_t maxg = tune_.max_grid < (uint32_t)MAXG ? tune_.max_grid : (uint32_t)MAXG;
    if (grid > maxg) grid = maxg;
    if (grid < 1) grid = 1;
    while (grid & (grid - 1)) grid &= grid - 1;
    return grid;
  }

  // Rounds 0-2 from the input (fused01_kernel, fold2_kernel) when a whole
  // stage starts with three pass rounds and round 2 is not the last.
  bool fused_first_rounds(uint32_t first, uint32_t last) const {
    if (!tune_.lookahead || first != 0 || last < 3 || n_ < 4) return false;
    const size_t pairs2 = (size_t)1 << (n_ - 3);
    return pairs2 > tune_.tail_pairs && pairs2 > tune_.chain_pairs;
  }

  int enqueue(uint32_t first, uint32_t last, hipStream_t s) override {
    if (refold_ && first < refold_) {
      // Sqrt-space rounds: refold the untouched input over the bound prefix;
      // the last one leaves its round table for the dense rounds after it.
      const uint32_t stop = last < refold_ ? last : refold_;
      const uint32_t maxg = tune_.max_grid < (uint32_t)MAXG ? tune_.max_grid : (uint32_t)MAXG;
      for (uint32_t r = first; r < stop; ++r) {
        const size_t pairs = (size_t)1 << (n_ - r - 1);
        if (r == 0) {
          // Nothing is bound yet: round 0 is the dense round 0.
          size_t grid = pass_grid(pairs);
          hipEvent_t ev = begin(s);
          pass_kernel<FT, FT, FC, K, EQ, false><<<(uint32_t)grid + 1, PASS_THREADS, 0, s>>>(
              args_, args_.in, args_.in_stride, 1u << n_, 0, last == 1 ? 1u : 0u);
          ZS_HIP(hipGetLastError());
          if (ev) mark(s, 0, KREC_PASS, (uint32_t)grid, pass_bytes(0), ev);
          continue;
        }
        const uint32_t ps = refold_pshift(pairs, maxg);
        size_t grid = pairs >> ps;
        if (grid > maxg) grid = maxg;
        const bool write = r + 1 == refold_ && refold_ >= 2;
        hipEvent_t ev = begin(s);
        refold_pass_kernel<FT, FC, K><<<(uint32_t)grid + 1, PASS_THREADS, 0, s>>>(
            args_, r, ps, write ? 1u : 0u, r + 1 == last ? 1u : 0u);
        ZS_HIP(hipGetLastError());
        if (ev) {
          uint64_t bytes = (uint64_t)K * ((uint64_t)1 << n_) * sizeof(TT);
          if (write) bytes += (uint64_t)K * 2 * pairs * sizeof(TC);
          mark(s, r, KREC_REFOLD, (uint32_t)grid, bytes, ev);
        }
      }
      if (stop == last) return ZS_OK;
      first = stop;
    }
    bool cont = false;  // the previous launch left this round's coefficients
    uint32_t r0 = first;
    if (fused_first_rounds(first, last)) {
      const size_t N = (size_t)1 << n_;
      const uint64_t in_bytes = (uint64_t)K * N * sizeof(TT);
      // A thread unit of fused01_kernel is a 16-byte vector of groups.
      size_t grid = pass_grid(N / 4 / (16 / sizeof(TT)));
      hipEvent_t ev = begin(s);
      fused01_kernel<FT, FC, K, EQ><<<(uint32_t)grid + 1, PASS_THREADS, 0, s>>>(args_);
      ZS_HIP(hipGetLastError());
      if (ev) mark(s, 0, KREC_FUSED01, (uint32_t)grid, in_bytes, ev);
      grid = pass_grid(N / 8);
      ev = begin(s);
      fold2_kernel<FT, FC, K, EQ>
          <<<(uint32_t)grid + 1, PASS_THREADS, 0, s>>>(args_, last == 3 ? 1u : 0u);
      ZS_HIP(hipGetLastError());
      if (ev) mark(s, 2, KREC_FOLD2, (uint32_t)grid, in_bytes + (uint64_t)K * (N / 4) * sizeof(TC), ev);
      r0 = 3;
    }
    for (uint32_t r = r0; r < last; ++r) {
      const size_t pairs = (size_t)1 << (n_ - r - 1);
      if (pairs <= tune_.tail_pairs) {
        hipEvent_t ev = begin(s);
        const bool la = tune_.lookahead && r >= 2 && last == n_ && n_ - r >= 2;
        if (la)
          tail_la_kernel<FT, FC, K, EQ><<<1, LA_THREADS, 0, s>>>(args_, r, cont ? 1u : 0u);
        else
          tail_kernel<FT, FC, K, EQ><<<1, TAIL_THREADS, 0, s>>>(args_, r, last);
        ZS_HIP(hipGetLastError());
        if (ev) {
          uint64_t b = 0;
          for (uint32_t q = r; q < last; ++q) b += pass_bytes(q);
          mark(s, r, la ? KREC_TAIL_LA : KREC_TAIL, 1, b, ev);
        }
        return ZS_OK;
      }
      if (r >= 2 && pairs <= tune_.chain_pairs) {
        // Chained: rounds r .. (tail or last) in one persistent launch.
        uint32_t end = r;
        while (end < last && ((size_t)1 << (n_ - end - 1)) > tune_.tail_pairs) ++end;
        const bool la = tune_.lookahead && end - r >= 2;
        int rc = la ? enqueue_chain_la(r, end, last, cont, s) : enqueue_chain(r, end, last, s);
        if (rc) return rc;
        // A look-ahead chain that stops before the last round leaves the next
        // round's coefficients for the tail.
        cont = la && end < n_;
        r = end - 1;
        continue;
      }
      size_t per_block = (size_t)PASS_THREADS * tune_.pairs_per_thread;
      size_t grid = (pairs + per_block - 1) / per_block;
      uint32_t maxg = tune_.max_grid < (uint32_t)MAXG ? tune_.max_grid : (uint32_t)MAXG;
      if (grid > maxg) grid = maxg;
      if (grid < 1) grid = 1;
      // A power-of-two grid keeps each thread's eq_r index fixed across its pairs.
      while (grid & (grid - 1)) grid &= grid - 1;
      const uint32_t L = 1u << (n_ - r);
      const uint32_t need_claim = r + 1 == last ? 1u : 0u;  // the claim is read after `last`
      hipEvent_t ev = begin(s);
      // Block 0 of a pass kernel finishes the round; grid blocks stream it.
      const uint32_t blocks = (uint32_t)grid + 1;
      if (r >= 2 && chain_la_starts(r + 1, last)) {
        // The round before a look-ahead chain also sums the chain's first
        // coefficients, so that the chain hashes its first message at once.
        pass_la_kernel<FT, FC, K, EQ><<<blocks, PASS_THREADS, 0, s>>>(args_, r);
        ZS_HIP(hipGetLastError());
        if (ev) mark(s, r, KREC_PASS_LA, (uint32_t)grid, pass_bytes(r), ev);
        cont = true;
        continue;
      }
      if (r == 0) {
        pass_kernel<FT, FT, FC, K, EQ, false>
            <<<blocks, PASS_THREADS, 0, s>>>(args_, args_.in, args_.in_stride, L, r, need_claim);
      } else if (r == 1) {
        pass_kernel<FT, FT, FC, K, EQ, true>
            <<<blocks, PASS_THREADS, 0, s>>>(args_, args_.in, args_.in_stride, L, r, need_claim);
      } else {
        pass_kernel<FT, FC, FC, K, EQ, true>
            <<<blocks, PASS_THREADS, 0, s>>>(args_, args_.work, args_.work_stride, L, r,
                                             need_claim);
      }
      ZS_HIP(hipGetLastError());
      if (ev) mark(s, r, KREC_PASS, (uint32_t)grid, pass_bytes(r), ev);
    }
    return ZS_OK;
  }

  // Rounds [first, end) in one launch of chain_kernel: one block per
  // PASS_THREADS * pairs_per_thread pairs of the first round, a power of two,
  // at most max_grid, chain_grid and half the blocks the device holds at
  // once, so that every block is resident even beside other work (blocks wait
  // for one another; their waits are bounded, so a violation faults the proof
  // instead of hanging the device).
  int enqueue_chain(uint32_t first, uint32_t end, uint32_t last, hipStream_t s) {
    static int resident = 0;  // per instantiation: blocks of this kernel the device holds
    auto kern = chain_kernel<FT, FC, K, EQ>;
    if (resident == 0) {
      int per_cu = 0, dev = 0, cus = 0;
      ZS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, PASS_THREADS, 0));
      ZS_HIP(hipGetDevice(&dev));
      ZS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      resident = per_cu * cus;
      if (resident < 1) return fail(ZS_EHIP, "chained rounds: kernel does not fit on the device");
    }
    const size_t pairs = (size_t)1 << (n_ - first - 1);
    const size_t per_block = (size_t)PASS_THREADS * tune_.pairs_per_thread;
    size_t grid = (pairs + per_block - 1) / per_block;
    uint32_t maxg = tune_.max_grid < (uint32_t)MAXG ? tune_.max_grid : (uint32_t)MAXG;
    if (grid > maxg) grid = maxg;
    if (grid > (size_t)resident / 2) grid = (size_t)resident / 2;
    if (grid > tune_.chain_grid) grid = tune_.chain_grid;
    if (grid < 1) grid = 1;
    while (grid & (grid - 1)) grid &= grid - 1;
    hipEvent_t ev = begin(s);
    kern<<<(uint32_t)grid, PASS_THREADS, 0, s>>>(args_, first, end, last);
    ZS_HIP(hipGetLastError());
    if (ev) {
      uint64_t b = 0;
      for (uint32_t q = first; q < end; ++q) b += pass_bytes(q);
      mark(s, first, KREC_CHAIN, (uint32_t)grid, b, ev);
    }
    return ZS_OK;
  }

  // Rounds [first, end) (at least two) in one launch of chain_la_kernel:
  // block 0 finishes, and blocks 1..grid, sized as in enqueue_chain from the
  // first period's pairs (round first+1's), stream the periods. `cont`: the
  // previous launch (pass_la_kernel) left round first's coefficients.
  int enqueue_chain_la(uint32_t first, uint32_t end, uint32_t last, bool cont, hipStream_t s) {
    static int resident = 0;  // per instantiation: blocks of this kernel the device holds
    auto kern = chain_la_kernel<FT, FC, K, EQ>;
    if (resident == 0) {
      int per_cu = 0, dev = 0, cus = 0;
      ZS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, PASS_THREADS, 0));
      ZS_HIP(hipGetDevice(&dev));
      ZS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      resident = per_cu * cus;
      if (resident < 2) return fail(ZS_EHIP, "chained rounds: kernel does not fit on the device");
    }
    const size_t units = (size_t)1 << (n_ - first - 2);
    const size_t per_block = (size_t)PASS_THREADS * tune_.pairs_per_thread;
    size_t grid = (units + per_block - 1) / per_block;
    uint32_t maxg = tune_.max_grid < (uint32_t)MAXG ? tune_.max_grid : (uint32_t)MAXG;
    if (grid > maxg) grid = maxg;
    if (grid > (size_t)resident / 2) grid = (size_t)resident / 2;
    if (grid > tune_.chain_grid) grid = tune_.chain_grid;
    if (grid < 1) grid = 1;
    while (grid & (grid - 1)) grid &= grid - 1;
    hipEvent_t ev = begin(s);
    kern<<<(uint32_t)grid + 1, PASS_THREADS, 0, s>>>(args_, first, end, last, cont ? 1u : 0u);
    ZS_HIP(hipGetLastError());
    if (ev) {
      uint64_t b = 0;
      for (uint32_t q = first; q < end; ++q) b += pass_bytes(q);
      mark(s, first, KREC_CHAIN_LA, (uint32_t)grid, b, ev);
    }
    return ZS_OK;
  }

  int download(hipStream_t s) override {
    ZS_HIP(hipMemcpyAsync(host_, block_, out_bytes_, hipMemcpyDeviceToHost, s));
    return ZS_OK;
  }

  int collect(zs_proof* out, zs_transcript* t) override {
    const uint8_t* h = (const uint8_t*)host_;
    const Ctl<FC>& c = *(const Ctl<FC>*)(h + o_ctl_);
    if (c.fault) return fail(ZS_EHIP, "chained rounds: a device wait reached its limit");
    if (out->msgs) copy_msg_rows(out->msgs, h + o_msgs_, n_, NPTS, mlen, sizeof(TC));
    if (out->point) memcpy(out->point, h + o_point_, (size_t)n_ * sizeof(TC));
    if (out->value) memcpy(out->value, &c.claim, sizeof(TC));
    if (out->final_tables) memcpy(out->final_tables, h + o_final_, (K + (EQ ? 1 : 0)) * sizeof(TC));
    if (t) memcpy(&t->sha, &c.tr, sizeof(TrState));
    return ZS_OK;
  }

  int upload_transcript(const zs_transcript& t, hipStream_t s) override {
    // The transcript prefix of the control block: state, word size, limbs.
    Ctl<FC>& c = *(Ctl<FC>*)((uint8_t*)host_ + o_ctl_);
    memcpy(&c.tr, &t.sha, sizeof(TrState));
    c.tr_kind = t.kind;
    c.item_bytes = t.item_bytes;
    c.limbs = (uint32_t)(sizeof(TC) / t.item_bytes);
    static_assert(offsetof(Ctl<FC>, limbs) + sizeof(uint32_t) == offsetof(Ctl<FC>, counter),
                  "control block prefix");
    ZS_HIP(hipMemcpyAsync(args_.ctl, &c, offsetof(Ctl<FC>, counter), hipMemcpyHostToDevice, s));
    return ZS_OK;
  }

  int download_round(uint32_t, hipStream_t s) override { return download(s); }

  int collect_round(uint32_t round, void* msg, void* chal, void* claim, zs_transcript* t) override {
    const uint8_t* h = (const uint8_t*)host_;
    const Ctl<FC>& c = *(const Ctl<FC>*)(h + o_ctl_);
    if (c.fault) return fail(ZS_EHIP, "chained rounds: a device wait reached its limit");
    if (msg) memcpy(msg, h + o_msgs_ + (size_t)round * NPTS * sizeof(TC), mlen * sizeof(TC));
    if (chal) memcpy(chal, h + o_point_ + (size_t)round * sizeof(TC), sizeof(TC));
    if (claim) memcpy(claim, &c.claim, sizeof(TC));
    if (t) memcpy(&t->sha, &c.tr, sizeof(TrState));
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
    if (refold_ && bound < refold_)
      return fail(ZS_EINVAL, "a sqrt-space state holds no round tables before round " +
                                 std::to_string(refold_));
    // The device holds the previous round's tables (the input after round
    // 0): fold a copy at the last challenge.
    const TC r = ((const Ctl<FC>*)((const uint8_t*)host_ + o_ctl_))->r;
    std::vector<TC> prev((size_t)K * 2 * len);
    if (bound == 1) {
      std::vector<TT> tmp((size_t)K * 2 * len);
      for (uint32_t q = 0; q < K; ++q)
        ZS_HIP(hipMemcpyAsync(tmp.data() + (size_t)q * 2 * len, args_.in + (size_t)q * args_.in_stride,
                              2 * len * sizeof(TT), hipMemcpyDeviceToHost, s));
      ZS_HIP(hipStreamSynchronize(s));
      for (size_t i = 0; i < tmp.size(); ++i) prev[i] = Lift<FT, FC>::go(tmp[i]);
    } else {
      for (uint32_t q = 0; q < K; ++q)
        ZS_HIP(hipMemcpyAsync(prev.data() + (size_t)q * 2 * len,
                              args_.work + (size_t)q * args_.work_stride, 2 * len * sizeof(TC),
                              hipMemcpyDeviceToHost, s));
      ZS_HIP(hipStreamSynchronize(s));
    }
    TC* o = (TC*)out;
    for (uint32_t q = 0; q < K; ++q)
      for (size_t j = 0; j < len; ++j) {
        const TC lo = prev[(size_t)q * 2 * len + j], hi = prev[(size_t)q * 2 * len + len + j];
        o[(size_t)q * len + j] = FC::add(lo, FC::mul(r, FC::sub(hi, lo)));
      }
    return ZS_OK;
  }

  void free_buffers() override {
    if (!external_ && block_ && ctx_) ctx_->pool->release(block_);
    if (!external_ && host_ && ctx_) ctx_->host_pool->release(host_);
    block_ = nullptr;
    host_ = nullptr;
  }

 private:
  zs_ctx* ctx_ = nullptr;
  uint32_t n_ = 0;
  uint32_t refold_ = 0;  // sqrt-space rounds before the dense ones
  TuneParams tune_;
  bool external_ = false;  // block_ and host_ belong to the caller
  void* block_ = nullptr;
  void* host_ = nullptr;
  size_t o_ctl_ = 0, o_msgs_ = 0, o_point_ = 0, o_final_ = 0, out_bytes_ = 0, io_bytes_ = 0;
  StageArgs<FT, FC> args_{};
};

// The InstanceImpl of K factor tables for a message of npts values (K + 1, or K + 2 with the eq factor).
template <class FT, class FC, int K>
Instance* make_typed_k(uint32_t npts, bool eq) {
  if (eq && npts == (uint32_t)(K + 2)) return new InstanceImpl<FT, FC, K, K + 2, true>();
  if (!eq && npts == (uint32_t)(K + 1)) return new InstanceImpl<FT, FC, K, K + 1, false>();
  return nullptr;
}

}  // namespace zs
