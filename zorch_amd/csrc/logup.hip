// The LogUp-GKR layer sumcheck on the device (reference
// zorch/logup_gkr/prover.py: logup_combine :56-68, LogupSumcheckRound
// :232-300, GkrLayerRound :344-378). The state is the reference's five
// stacked tables [eq, n0, d1, n1, d0], folded together each round (MSB
// first); round r's message is the round polynomial at t = 0..3 of
//   sum_x eq (lam (n0 d1 + n1 d0) + d0 d1),
// every factor at (t, x). One launch a round: blocks 1..G fold the previous
// round's tables at its challenge and sum the four node values of their
// pairs into the slot rows; block 0 finishes the round (finish_round: the
// message, the Fiat-Shamir hop, the challenge, the claim reduction and, after
// the last round, the five folded values). Tables and challenges share one
// field (the reference's layers hold extension-field fractions).
#include "instance.h"

namespace zs {

constexpr int LG_TABLES = 5;  // [eq, n0, d1, n1, d0]
constexpr int LG_NA = 4;      // message nodes t = 0..3 (degree 3)
// Rounds with at most this many pairs run in one block (logup_tail_kernel).
constexpr size_t LG_TAIL_PAIRS = 512;

// eq (lam (n0 d1 + n1 d0) + d0 d1) on v = [eq, n0, d1, n1, d0].
template <class F>
__device__ __forceinline__ typename F::T logup_combine(typename F::T lam,
                                                       const typename F::T (&v)[LG_TABLES]) {
  const typename F::T cross = F::add(F::mul(v[1], v[2]), F::mul(v[3], v[4]));
  return F::mul(v[0], F::add(F::mul(lam, cross), F::mul(v[4], v[2])));
}

// Pairs start, start + stride, ... of a round of L entries a table: fold
// (FOLD: `src` holds the previous round's tables, 2L entries each, folded at
// r into `work`; pair j reads and writes only positions j, j + L/2 of this
// round and j + L, j + 3L/2 of the previous, so the fold runs in place) and
// add the summand at t = 0..3 to acc.
template <class F, bool FOLD>
__device__ __forceinline__ void logup_pairs(const StageArgs<F, F>& a, const typename F::T* src,
                                            size_t src_stride, uint32_t L, typename F::T lam,
                                            typename F::T r, size_t start, size_t stride,
                                            typename F::T (&acc)[LG_NA]) {
  using T = typename F::T;
  const size_t half = L >> 1;
  for (size_t j = start; j < half; j += stride) {
    T lo[LG_TABLES], df[LG_TABLES];
#pragma unroll
    for (int q = 0; q < LG_TABLES; ++q) {
      const T* s = src + (size_t)q * src_stride;
      T x0, x1;
      if constexpr (FOLD) {
        const T a0 = s[j], a1 = s[j + L], b0 = s[j + half], b1 = s[j + half + L];
        x0 = F::add(a0, F::mul(r, F::sub(a1, a0)));
        x1 = F::add(b0, F::mul(r, F::sub(b1, b0)));
        T* d = a.work + (size_t)q * a.work_stride;
        d[j] = x0;
        d[j + half] = x1;
      } else {
        x0 = s[j];
        x1 = s[j + half];
      }
      lo[q] = x0;
      df[q] = F::sub(x1, x0);
    }
#pragma unroll
    for (int t = 0; t < LG_NA; ++t) {
      acc[t] = F::add(acc[t], logup_combine<F>(lam, lo));
      if (t + 1 < LG_NA) {
#pragma unroll
        for (int q = 0; q < LG_TABLES; ++q) lo[q] = F::add(lo[q], df[q]);
      }
    }
  }
}

// A small round in one block (4 x pairs <= block threads): thread 4j + t
// folds pair j (only t = 0 writes the folded entries) and adds the summand
// at node t alone, so a pair's four combines run side by side.
template <class F, bool FOLD>
__device__ __forceinline__ void logup_pairs_split(const StageArgs<F, F>& a,
                                                  const typename F::T* src, size_t src_stride,
                                                  uint32_t L, typename F::T lam, typename F::T r,
                                                  typename F::T (&acc)[LG_NA]) {
  using T = typename F::T;
  const size_t half = L >> 1;
  const size_t j = threadIdx.x >> 2;
  const uint32_t t = threadIdx.x & 3u;
  if (j >= half) return;
  T v[LG_TABLES];
#pragma unroll
  for (int q = 0; q < LG_TABLES; ++q) {
    const T* s = src + (size_t)q * src_stride;
    T x0, x1;
    if constexpr (FOLD) {
      const T a0 = s[j], a1 = s[j + L], b0 = s[j + half], b1 = s[j + half + L];
      x0 = F::add(a0, F::mul(r, F::sub(a1, a0)));
      x1 = F::add(b0, F::mul(r, F::sub(b1, b0)));
      if (t == 0) {
        T* d = a.work + (size_t)q * a.work_stride;
        d[j] = x0;
        d[j + half] = x1;
      }
    } else {
      x0 = s[j];
      x1 = s[j + half];
    }
    const T df = F::sub(x1, x0);
    T x = x0;
    for (uint32_t i = 0; i < t; ++i) x = F::add(x, df);
    v[q] = x;
  }
  const T c = logup_combine<F>(lam, v);
#pragma unroll
  for (int i = 0; i < LG_NA; ++i)
    if ((uint32_t)i == t) acc[i] = F::add(acc[i], c);
}

// Round `round` of a LogUp stage as 1 + nblk blocks. Blocks 1..nblk stream
// the pairs (logup_pairs), add their sums to the slot rows and arrive; block
// 0 warms the hop on a scratch copy of the transcript state (the finisher's
// code then sits in its compute unit's instruction cache), waits for the
// arrivals and finishes the round (finish_round: the message, the
// Fiat-Shamir hop, the challenge, the claim reduction and, after the last
// round, the five folded values).
template <class F, bool FOLD>
__global__ __launch_bounds__(PASS_THREADS) void logup_pass_kernel(StageArgs<F, F> a,
                                                                  const typename F::T* src,
                                                                  size_t src_stride, uint32_t L,
                                                                  uint32_t round,
                                                                  uint32_t need_claim) {
  using T = typename F::T;
  __shared__ PassLds<F, LG_NA> ps;
  RoundLds<F, LG_NA>& sh = ps.round;
  const uint32_t tid = threadIdx.x, nblk = gridDim.x - 1;
  Ctl<F>* ctl = a.ctl;
  if (blockIdx.x == 0) {
    if (tid < 64) {
      fin_load(sh.sm, ctl);
      const uint32_t* from = reinterpret_cast<const uint32_t*>(static_cast<HopLds*>(&sh.sm));
      uint32_t* to = reinterpret_cast<uint32_t*>(&ps.warm);
      for (uint32_t i = tid; i < sizeof(HopLds) / 4; i += 64u) to[i] = from[i];
      wave_sync();
      fs_hop<HOP_PASS>((ZS_LDS HopLds*)&ps.warm, (const ZS_LDS uint8_t*)sh.sm.msg,
                       (uint32_t)(LG_NA * sizeof(T)), ctl->item_bytes, ctl->limbs, ctl->tr_kind);
    }
    if (tid == 0) sh.is_last = chain_wait(&ctl->counter, nblk, &ctl->fault) ? 1u : 0u;
    __syncthreads();
    if (!sh.is_last) return;
    slot_rows_sum<F, LG_NA>(a.slots, sh);
    if (tid < 64) {
      // The message shape of a degree-3 product round without the eq factor
      // (the eq table is folded like the others).
      finish_round<F, F, LG_NA - 1, false, HOP_PASS>(a, round, need_claim != 0, sh.out, sh.sm,
                                                     a.work, a.in,
                                                     round == 0 ? a.in_stride : a.work_stride,
                                                     round == 0);
      fin_store(sh.sm, ctl);
    }
    if (tid == 0) __hip_atomic_store(&ctl->counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  const uint32_t blk = blockIdx.x - 1;
  T acc[LG_NA];
#pragma unroll
  for (int t = 0; t < LG_NA; ++t) acc[t] = F::zero();
  logup_pairs<F, FOLD>(a, src, src_stride, L, ctl->lam, ctl->r,
                       (size_t)blk * PASS_THREADS + tid, (size_t)nblk * PASS_THREADS, acc);
  block_reduce<F, LG_NA, PASS_THREADS>(acc, sh.lds, sh.out);
  slot_row_add<F, LG_NA>(a.slots, blk, sh.out);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (tid == 0) __hip_atomic_fetch_add(&ctl->counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Rounds first..last-1 of a LogUp stage in one block: each round streams its
// pairs over the block, reduces in LDS and wave 0 finishes it; the next round
// folds at the challenge the finisher left in LDS. For the rounds whose pairs
// are too few to fill the device, where a launch per round and a device-wide
// arrival cost more than the pairs.
template <class F>
__global__ __launch_bounds__(PASS_THREADS) void logup_tail_kernel(StageArgs<F, F> a,
                                                                  uint32_t first, uint32_t last) {
  using T = typename F::T;
  __shared__ RoundLds<F, LG_NA> sh;
  const uint32_t tid = threadIdx.x;
  Ctl<F>* ctl = a.ctl;
  if (tid < 64) fin_load(sh.sm, ctl);
  __syncthreads();
  const T lam = ctl->lam;
  for (uint32_t round = first; round < last; ++round) {
    const uint32_t L = 1u << (a.n - round);
    T acc[LG_NA];
#pragma unroll
    for (int t = 0; t < LG_NA; ++t) acc[t] = F::zero();
    const T r = sh.sm.r;
    const bool split = (size_t)(L >> 1) * 4 <= PASS_THREADS;
    if (round == 0) {
      if (split) logup_pairs_split<F, false>(a, a.in, a.in_stride, L, lam, r, acc);
      else logup_pairs<F, false>(a, a.in, a.in_stride, L, lam, r, tid, PASS_THREADS, acc);
    } else if (round == 1) {
      if (split) logup_pairs_split<F, true>(a, a.in, a.in_stride, L, lam, r, acc);
      else logup_pairs<F, true>(a, a.in, a.in_stride, L, lam, r, tid, PASS_THREADS, acc);
    } else {
      if (split) logup_pairs_split<F, true>(a, a.work, a.work_stride, L, lam, r, acc);
      else logup_pairs<F, true>(a, a.work, a.work_stride, L, lam, r, tid, PASS_THREADS, acc);
    }
    // The folded entries reach the next round's readers through the barriers
    // (one workgroup).
    block_reduce<F, LG_NA, PASS_THREADS>(acc, sh.lds, sh.out);
    if (tid < 64)
      finish_round<F, F, LG_NA - 1, false, HOP_INLINE>(a, round, round + 1 == last, sh.out, sh.sm,
                                                     a.work, a.in,
                                                     round == 0 ? a.in_stride : a.work_stride,
                                                     round == 0);
    __syncthreads();
  }
  if (tid < 64) fin_store(sh.sm, ctl);
}

// A LogUp stage (zs_prove, zs_state_*): buffers laid out as the product
// stages' (instance.h), five tables, messages of four values.
template <class F>
class LogupInstance final : public Instance {
  using T = typename F::T;

 public:
  ~LogupInstance() override { free_buffers(); }

  using Instance::setup;
  int setup(zs_ctx* ctx, const zs_prove_args& a, const zs_transcript& t, hipStream_t s,
            const TuneParams& tune) override {
    return setup_in(ctx, a, t, s, tune, nullptr, 0, nullptr, 0);
  }

  int setup_in(zs_ctx* ctx, const zs_prove_args& a, const zs_transcript& t, hipStream_t s,
               const TuneParams& tune, void* dev_mem, size_t, void*, size_t) override {
    if (dev_mem) return fail(ZS_EINVAL, "LogUp stages allocate their own buffers");
    ctx_ = ctx;
    n_ = a.rounds;
    rounds = n_;
    npts = LG_NA;
    k = LG_TABLES;
    chal_bytes = sizeof(T);
    tune_ = tune;
    const size_t N = (size_t)1 << n_;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off = align256(off + bytes); return o; };
    o_ctl_ = take(sizeof(Ctl<F>));
    o_msgs_ = take((size_t)n_ * LG_NA * sizeof(T));
    o_point_ = take((size_t)n_ * sizeof(T));
    o_final_ = take((size_t)LG_TABLES * sizeof(T));
    out_bytes_ = off;
    const size_t o_slots = take(((size_t)NSLOT * SLOTW + LA_SLOT_WORDS) * sizeof(uint64_t));
    io_bytes_ = off;
    const size_t work_stride = N / 2 > 0 ? N / 2 : 1;
    const size_t o_work = take((size_t)LG_TABLES * work_stride * sizeof(T));
    const bool stage_in = !a.tables_on_device;
    const size_t o_in = stage_in ? take((size_t)LG_TABLES * N * sizeof(T)) : 0;
    int rc;
    if ((rc = ctx->pool->alloc(off, &block_))) return rc;
    device_bytes = off;
    if ((rc = ctx->host_pool->alloc(io_bytes_, &host_))) return rc;
    uint8_t* b = (uint8_t*)block_;
    args_.in = stage_in ? (const T*)(b + o_in) : (const T*)a.tables;
    args_.in_stride = N;
    args_.work = (T*)(b + o_work);
    args_.work_stride = work_stride;
    args_.n = n_;
    args_.k = LG_TABLES;
    args_.ctl = (Ctl<F>*)(b + o_ctl_);
    args_.slots = (uint64_t*)(b + o_slots);
    args_.msgs = (T*)(b + o_msgs_);
    args_.point = b + o_point_;
    args_.final_out = (T*)(b + o_final_);
    args_.trace = ctx->trace;
    if (stage_in)
      ZS_HIP(hipMemcpyAsync((void*)args_.in, a.tables, (size_t)LG_TABLES * N * sizeof(T),
                            hipMemcpyHostToDevice, s));
    uint8_t* h = (uint8_t*)host_;
    memset(h, 0, io_bytes_);
    Ctl<F>& c = *(Ctl<F>*)(h + o_ctl_);
    init_ctl_consts<F, LG_NA - 1, LG_NA>(c);
    if ((rc = set_ctl_proof<F>(c, t, a, LG_NA))) return rc;
    c.lam = load_host_el<F>(a.lambda, 0);
    mlen = c.mlen;
    ZS_HIP(hipMemcpyAsync(b, h, io_bytes_, hipMemcpyHostToDevice, s));
    return ZS_OK;
  }

  int enqueue(uint32_t first, uint32_t last, hipStream_t s) override {
    for (uint32_t r = first; r < last; ++r) {
      const uint32_t L = 1u << (n_ - r);
      const size_t pairs = L / 2;
      if (pairs <= LG_TAIL_PAIRS) {
        // Every remaining round in one block.
        hipEvent_t ev = nullptr;
        if (ctx_->profile) {
          ev = ctx_->event();
          (void)hipEventRecord(ev, s);
        }
        logup_tail_kernel<F><<<1, PASS_THREADS, 0, s>>>(args_, r, last);
        ZS_HIP(hipGetLastError());
        if (ev) {
          hipEvent_t stop = ctx_->event();
          (void)hipEventRecord(stop, s);
          const uint64_t b = (uint64_t)LG_TABLES * 3 * L * sizeof(T);
          ctx_->records.push_back(KernelRecord{r, KREC_TAIL, 1, b, ev, stop});
        }
        break;
      }
      const size_t per_block = (size_t)PASS_THREADS * tune_.pairs_per_thread;
      size_t grid = (pairs + per_block - 1) / per_block;
      const uint32_t maxg = tune_.max_grid < (uint32_t)MAXG ? tune_.max_grid : (uint32_t)MAXG;
      if (grid > maxg) grid = maxg;
      if (grid < 1) grid = 1;
      const uint32_t blocks = (uint32_t)grid + 1;
      const uint32_t need_claim = r + 1 == last ? 1u : 0u;  // the claim is read after `last`
      hipEvent_t ev = nullptr;
      if (ctx_->profile) {
        ev = ctx_->event();
        (void)hipEventRecord(ev, s);
      }
      if (r == 0) {
        logup_pass_kernel<F, false>
            <<<blocks, PASS_THREADS, 0, s>>>(args_, args_.in, args_.in_stride, L, r, need_claim);
      } else if (r == 1) {
        logup_pass_kernel<F, true>
            <<<blocks, PASS_THREADS, 0, s>>>(args_, args_.in, args_.in_stride, L, r, need_claim);
      } else {
        logup_pass_kernel<F, true><<<blocks, PASS_THREADS, 0, s>>>(args_, args_.work,
                                                                    args_.work_stride, L, r,
                                                                    need_claim);
      }
      ZS_HIP(hipGetLastError());
      if (ev) {
        // Five tables: round 0 reads L entries each; later rounds read the
        // previous round's 2L and write L.
        const uint64_t w = sizeof(T);
        const uint64_t b = r == 0 ? (uint64_t)LG_TABLES * L * w : (uint64_t)LG_TABLES * 3 * L * w;
        hipEvent_t stop = ctx_->event();
        (void)hipEventRecord(stop, s);
        ctx_->records.push_back(KernelRecord{r, KREC_LOGUP, (uint32_t)grid, b, ev, stop});
      }
    }
    return ZS_OK;
  }

  int download(hipStream_t s) override {
    ZS_HIP(hipMemcpyAsync(host_, block_, out_bytes_, hipMemcpyDeviceToHost, s));
    return ZS_OK;
  }

  int collect(zs_proof* out, zs_transcript* t) override {
    const uint8_t* h = (const uint8_t*)host_;
    const Ctl<F>& c = *(const Ctl<F>*)(h + o_ctl_);
    if (c.fault) return fail(ZS_EHIP, "LogUp rounds: a device wait reached its limit");
    if (out->msgs) copy_msg_rows(out->msgs, h + o_msgs_, n_, LG_NA, mlen, sizeof(T));
    if (out->point) memcpy(out->point, h + o_point_, (size_t)n_ * sizeof(T));
    if (out->value) memcpy(out->value, &c.claim, sizeof(T));
    if (out->final_tables) memcpy(out->final_tables, h + o_final_, LG_TABLES * sizeof(T));
    if (t) memcpy(&t->sha, &c.tr, sizeof(TrState));
    return ZS_OK;
  }

  int upload_transcript(const zs_transcript& t, hipStream_t s) override {
    Ctl<F>& c = *(Ctl<F>*)((uint8_t*)host_ + o_ctl_);
    memcpy(&c.tr, &t.sha, sizeof(TrState));
    c.tr_kind = t.kind;
    c.item_bytes = t.item_bytes;
    c.limbs = (uint32_t)(sizeof(T) / t.item_bytes);
    ZS_HIP(hipMemcpyAsync(args_.ctl, &c, offsetof(Ctl<F>, counter), hipMemcpyHostToDevice, s));
    return ZS_OK;
  }

  int download_round(uint32_t, hipStream_t s) override { return download(s); }

  int collect_round(uint32_t round, void* msg, void* chal, void* claim, zs_transcript* t) override {
    const uint8_t* h = (const uint8_t*)host_;
    const Ctl<F>& c = *(const Ctl<F>*)(h + o_ctl_);
    if (c.fault) return fail(ZS_EHIP, "LogUp rounds: a device wait reached its limit");
    if (msg) memcpy(msg, h + o_msgs_ + (size_t)round * LG_NA * sizeof(T), mlen * sizeof(T));
    if (chal) memcpy(chal, h + o_point_ + (size_t)round * sizeof(T), sizeof(T));
    if (claim) memcpy(claim, &c.claim, sizeof(T));
    if (t) memcpy(&t->sha, &c.tr, sizeof(TrState));
    return ZS_OK;
  }

  // The five tables after `bound` rounds.
  int read_tables(uint32_t bound, void* out, size_t out_bytes, hipStream_t s) override {
    const size_t len = (size_t)1 << (n_ - bound);
    if (out_bytes < (size_t)LG_TABLES * len * sizeof(T))
      return fail(ZS_EINVAL, "output buffer too small");
    if (bound == n_) {
      ZS_HIP(hipMemcpyAsync(out, args_.final_out, LG_TABLES * sizeof(T), hipMemcpyDeviceToHost, s));
      ZS_HIP(hipStreamSynchronize(s));
      return ZS_OK;
    }
    // Round `bound`'s tables: the input, or the previous round's (the
    // device holds them) folded at the last challenge.
    const T* src = bound <= 1 ? args_.in : args_.work;
    const size_t stride = bound <= 1 ? args_.in_stride : args_.work_stride;
    const size_t have = bound == 0 ? len : 2 * len;
    std::vector<T> prev((size_t)LG_TABLES * have);
    for (uint32_t q = 0; q < (uint32_t)LG_TABLES; ++q)
      ZS_HIP(hipMemcpyAsync(prev.data() + (size_t)q * have, src + (size_t)q * stride,
                            have * sizeof(T), hipMemcpyDeviceToHost, s));
    ZS_HIP(hipStreamSynchronize(s));
    T* o = (T*)out;
    if (bound == 0) {
      memcpy(o, prev.data(), prev.size() * sizeof(T));
      return ZS_OK;
    }
    const T r = ((const Ctl<F>*)((const uint8_t*)host_ + o_ctl_))->r;
    for (uint32_t q = 0; q < (uint32_t)LG_TABLES; ++q)
      for (size_t j = 0; j < len; ++j) {
        const T lo = prev[(size_t)q * have + j], hi = prev[(size_t)q * have + len + j];
        o[(size_t)q * len + j] = F::add(lo, F::mul(r, F::sub(hi, lo)));
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
  zs_ctx* ctx_ = nullptr;
  uint32_t n_ = 0;
  TuneParams tune_;
  void* block_ = nullptr;
  void* host_ = nullptr;
  size_t o_ctl_ = 0, o_msgs_ = 0, o_point_ = 0, o_final_ = 0, out_bytes_ = 0, io_bytes_ = 0;
  StageArgs<F, F> args_{};
};

// The fractional-sum circuit on the device. Layers live in 5-row buffers
// laid out as a LogUp stage's tables [eq, n0, d1, n1, d0] (row stride =
// the layer's width), so a layer is proved where it was built: row 0 takes
// the eq table of the layer's point (logup_eq_kernel), rows 1..4 the
// children. Next layer, pair (2j, 2j + 1) (reference circuit.py:100-140):
// n0' = n0 d1 + n1 d0 and d0' = d0 d1 at 2j, n1', d1' the same at 2j + 1.
template <class F>
__global__ __launch_bounds__(PASS_THREADS) void logup_transition_kernel(
    const typename F::T* in, size_t in_stride, typename F::T* out, size_t out_stride,
    size_t half) {
  using T = typename F::T;
  for (size_t j = (size_t)blockIdx.x * PASS_THREADS + threadIdx.x; j < half;
       j += (size_t)gridDim.x * PASS_THREADS) {
    T n0[2], d1[2], n1[2], d0[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      n0[h] = in[1 * in_stride + 2 * j + h];
      d1[h] = in[2 * in_stride + 2 * j + h];
      n1[h] = in[3 * in_stride + 2 * j + h];
      d0[h] = in[4 * in_stride + 2 * j + h];
    }
    out[1 * out_stride + j] = F::add(F::mul(n0[0], d1[0]), F::mul(n1[0], d0[0]));  // n0'
    out[2 * out_stride + j] = F::mul(d0[1], d1[1]);                                // d1'
    out[3 * out_stride + j] = F::add(F::mul(n0[1], d1[1]), F::mul(n1[1], d0[1]));  // n1'
    out[4 * out_stride + j] = F::mul(d0[0], d1[0]);                                // d0'
  }
}

// eq(z, x) for x in {0,1}^n, z_0 on the index MSB (reference
// zorch/poly/eq.py expand_eq_to_hypercube): each entry its product over
// the coordinates.
template <class F>
__global__ __launch_bounds__(PASS_THREADS) void logup_eq_kernel(const typename F::T* z, uint32_t n,
                                                                typename F::T* out) {
  using T = typename F::T;
  __shared__ T zs[64];
  if (threadIdx.x < n) zs[threadIdx.x] = z[threadIdx.x];
  __syncthreads();
  const size_t N = (size_t)1 << n;
  for (size_t x = (size_t)blockIdx.x * PASS_THREADS + threadIdx.x; x < N;
       x += (size_t)gridDim.x * PASS_THREADS) {
    T v = F::one();
    for (uint32_t i = 0; i < n; ++i)
      v = F::mul(v, ((x >> (n - 1 - i)) & 1) ? zs[i] : F::sub(F::one(), zs[i]));
    out[x] = v;
  }
}

static uint32_t logup_grid(size_t work) {
  size_t g = (work + PASS_THREADS - 1) / PASS_THREADS;
  return (uint32_t)(g < 1 ? 1 : g > 4096 ? 4096 : g);
}

template <class F>
static int logup_transition_typed(zs_ctx* ctx, const void* layer, size_t width, void* next) {
  using T = typename F::T;
  logup_transition_kernel<F><<<logup_grid(width / 2), PASS_THREADS, 0, ctx->stream>>>(
      (const T*)layer, width, (T*)next, width / 2, width / 2);
  ZS_HIP(hipGetLastError());
  ZS_HIP(hipStreamSynchronize(ctx->stream));
  return ZS_OK;
}

template <class F>
static int logup_eq_typed(zs_ctx* ctx, const void* point, uint32_t n, void* out) {
  using T = typename F::T;
  void* zd = nullptr;
  int rc;
  if ((rc = ctx->pool->alloc(64 * sizeof(T), &zd))) return rc;
  std::vector<T> z(n);
  for (uint32_t i = 0; i < n; ++i) z[i] = load_host_el<F>(point, i);
  hipError_t e = hipMemcpyAsync(zd, z.data(), n * sizeof(T), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    logup_eq_kernel<F><<<logup_grid((size_t)1 << n), PASS_THREADS, 0, ctx->stream>>>(
        (const T*)zd, n, (T*)out);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx->pool->release(zd);
  if (e != hipSuccess) return hip_fail(e, "logup eq table");
  return ZS_OK;
}

int logup_transition(zs_ctx* ctx, uint32_t field, const void* layer, size_t width, void* next) {
  switch (field) {
    case ZS_F_BABYBEAR: return logup_transition_typed<BB>(ctx, layer, width, next);
    case ZS_F_BABYBEAR4: return logup_transition_typed<BB4>(ctx, layer, width, next);
    case ZS_F_GOLDILOCKS: return logup_transition_typed<GL>(ctx, layer, width, next);
    default: return fail(ZS_EINVAL, "unknown field");
  }
}

int logup_eq(zs_ctx* ctx, uint32_t field, const void* point, uint32_t n, void* out) {
  switch (field) {
    case ZS_F_BABYBEAR: return logup_eq_typed<BB>(ctx, point, n, out);
    case ZS_F_BABYBEAR4: return logup_eq_typed<BB4>(ctx, point, n, out);
    case ZS_F_GOLDILOCKS: return logup_eq_typed<GL>(ctx, point, n, out);
    default: return fail(ZS_EINVAL, "unknown field");
  }
}

Instance* make_logup_instance(uint32_t field) {
  switch (field) {
    case ZS_F_BABYBEAR: return new LogupInstance<BB>();
    case ZS_F_BABYBEAR4: return new LogupInstance<BB4>();
    case ZS_F_GOLDILOCKS: return new LogupInstance<GL>();
    default: return nullptr;
  }
}

}  // namespace zs
