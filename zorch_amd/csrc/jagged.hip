// The jagged LogUp-GKR layer sumcheck on the device (reference
// zorch/logup_gkr/jagged_prover.py:272-448, _jagged_rounds.py,
// zorch/sumcheck/jagged/rounds.py:36-85, schedule.py:212-257).
//
// A layer holds row_counts[s] rows of segment s (batch-major) of its 2^nrv
// virtual rows; the rest are the neutral fraction (n = 0, d = 1). The row
// variables bind LSB first. Round k's state holds counts_k[s] =
// ceil(rc[s] / 2^k) rows of segment s (exact layout, segments back to back);
// its pair j of segment s is rows (2j, 2j + 1), the odd one neutral past the
// count. One launch a row round: blocks 1..G take pairs (segment found by a
// binary search over the round's cumulative pair counts), fold round k-1's
// rows 4j..4j+3 at r_{k-1} into rows 2j, 2j+1 of round k (written to the
// other state buffer) and sum, with the pair's eq weight
// W = E_{k+1}[j] eq_int[s] (E_t the eq table of the first t row
// coordinates; the reference's folded eq_row is E times the bound factors,
// which the finisher applies as pad_adj):
//   ez += W B(even), eh += W B(even + odd), es += W,
//   B(n0, n1, d0, d1) = lam (n0 d1 + n1 d0) + d0 d1.
// Block 0 is the finisher: s(0) = (pad (1 - z) ez + (pad - pad es)(1 - z)) eq_adj,
// s(1/2) = (pad eh + 4 (pad - pad es)) / 8 eq_adj, s(1) = claim - s(0), the
// coefficients through (0, 1, 1/2, b) by a host-built inverse Vandermonde,
// the Fiat-Shamir hop, claim = s(r), pad *= eq(r, z). After the row rounds
// one block runs the batch rounds: the last row fold leaves one element a
// segment (2^niv), which folds densely under eq_int, then the final fold
// gives the four openings.
#include "instance.h"

namespace zs {

constexpr int JG_NS = 3;  // sums a round: ez, eh, es
// Row rounds with at most this many pairs run in the one-block kernel.
constexpr uint32_t JG_TAIL_PAIRS = 2048;
// Layers of at most this many segments search a round's cumulative pair
// counts in LDS (staged once per block and round) instead of global memory.
constexpr uint32_t JG_LDS_SEGS = 4096;

// Row round k, segment s.
struct JagSeg {
  uint32_t cum_pairs;   // pairs of segments 0..s (inclusive)
  uint32_t base;        // first row of segment s in round k's state
  uint32_t count;       // rows of segment s in round k
  uint32_t prev_base;   // the same in round k - 1's state
  uint32_t prev_count;
  uint32_t pad_[3];
};

// Round constants: z (the round's eq coordinate), 1 - z and the map from
// (s(0), s(1), s(1/2)) to the four coefficients (the fourth node, b, has
// value zero).
template <class F>
struct JagRound {
  typename F::T z, omz;
  typename F::T m[4][3];
};

template <class F>
struct JagCtl {
  Ctl<F> c;              // transcript, claim, challenge, counter, fault, lam
  typename F::T pad;     // eq mass of the bound variables of the current phase
  typename F::T eq_adj;  // the row phase's eq mass once the row variables are bound
  typename F::T inv8, four;
};

template <class F>
struct JagArgs {
  const typename F::T* in;  // round 0's planes (the input)
  size_t in_stride;
  typename F::T* state[2];  // round k's planes: state[k & 1] (k >= 1)
  size_t state_stride;
  typename F::T* dense[2];  // the batch phase's planes (2^niv)
  size_t dense_stride;
  const JagSeg* segs;       // nrv x nseg
  const JagRound<F>* rc;    // niv + nrv
  const typename F::T* erow;  // E_t of the row coordinates at offset 2^t - 1
  const typename F::T* eint;  // the same of the batch coordinates
  const typename F::T* zrow;  // the row coordinates (nrv)
  const typename F::T* zint;  // the batch coordinates (niv)
  JagCtl<F>* ctl;
  uint64_t* slots;
  typename F::T* msgs;        // rounds x 4
  uint8_t* point;             // rounds x sizeof(T) raw squeezed bytes
  typename F::T* openings;    // 4
  uint32_t nseg, nrv, niv;
};

template <class F>
__device__ __forceinline__ typename F::T jag_bracket(typename F::T lam, typename F::T n0,
                                                     typename F::T n1, typename F::T d0,
                                                     typename F::T d1) {
  return F::add(F::mul(lam, F::add(F::mul(n0, d1), F::mul(n1, d0))), F::mul(d0, d1));
}

// E_t of z[0..t) (t = 0..tmax) at offsets 2^t - 1: E_t[j] = prod_i (bit
// t-1-i of j ? z_i : 1 - z_i), z_0 on the index MSB.
template <class F>
__global__ __launch_bounds__(PASS_THREADS) void jagged_eq_kernel(const typename F::T* z,
                                                                 uint32_t tmax,
                                                                 typename F::T* fam) {
  using T = typename F::T;
  const size_t total = ((size_t)2 << tmax) - 1;
  for (size_t idx = (size_t)blockIdx.x * PASS_THREADS + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * PASS_THREADS) {
    const uint32_t t = 63u - (uint32_t)__builtin_clzll((unsigned long long)(idx + 1));
    const size_t j = idx + 1 - ((size_t)1 << t);
    T v = F::one();
    for (uint32_t i = 0; i < t; ++i) {
      const T zi = z[i];
      v = F::mul(v, ((j >> (t - 1 - i)) & 1) ? zi : F::sub(F::one(), zi));
    }
    fam[idx] = v;
  }
}

// The finisher of round `round` (wave 0): message from the sums `s` (LDS),
// hop, challenge, claim and eq masses. The proof state is in `sm`, `pad` and
// `eq_adj` (every lane holds the same values).
template <class F, int HOP>
__device__ __forceinline__ void jag_finish(const JagArgs<F>& a, uint32_t round,
                                           const typename F::T* s, FinishLds<F>& sm,
                                           typename F::T& pad, typename F::T& eq_adj) {
  using T = typename F::T;
  constexpr int EB = sizeof(T);
  const uint32_t lane = threadIdx.x & 63;
  const JagRound<F>& R = a.rc[round];
  const T omz = R.omz;
  const T ez = F::mul(F::mul(pad, omz), s[0]);
  const T eh = F::mul(pad, s[1]);
  const T es = F::mul(pad, s[2]);
  const T corr = F::sub(pad, es);
  const T s0 = F::mul(F::add(ez, F::mul(corr, omz)), eq_adj);
  const T sh = F::mul(F::mul(F::add(eh, F::mul(corr, a.ctl->four)), a.ctl->inv8), eq_adj);
  const T s1 = F::sub(sm.claim, s0);
  if (lane < 4u) {
    const T c = F::add(F::add(F::mul(R.m[lane][0], s0), F::mul(R.m[lane][1], s1)),
                       F::mul(R.m[lane][2], sh));
    a.msgs[(size_t)round * 4 + lane] = c;
    sm.msg[lane] = c;
  }
  wave_sync();
  const Ctl<F>* ctl = &a.ctl->c;
  if constexpr (HOP == HOP_INLINE)
    fs_hop_body(sm, reinterpret_cast<const uint8_t*>(sm.msg), 4u * (uint32_t)EB, ctl->item_bytes,
                ctl->limbs, ctl->tr_kind);
  else
    fs_hop<HOP>((ZS_LDS HopLds*)static_cast<HopLds*>(&sm), (const ZS_LDS uint8_t*)sm.msg,
                4u * (uint32_t)EB, ctl->item_bytes, ctl->limbs, ctl->tr_kind);
  if (lane < (uint32_t)EB) a.point[(size_t)round * EB + lane] = sm.out[lane];
  const T r = el_from_words<F>(reinterpret_cast<const uint32_t*>(sm.out));
  T claim = sm.msg[3];
  for (int i = 2; i >= 0; --i) claim = F::add(F::mul(claim, r), sm.msg[i]);
  // eq(r, z) = r z + (1 - r)(1 - z)
  const T fac = F::add(F::mul(r, R.z), F::mul(F::sub(F::one(), r), omz));
  pad = F::mul(pad, fac);
  if (round + 1 == a.nrv) {
    eq_adj = pad;
    pad = F::one();
  }
  wave_sync();
  if (lane == 0) {
    sm.r = r;
    sm.claim = claim;
  }
  wave_sync();
}

// The neutral fraction's plane q (n0, n1, d0, d1).
template <class F>
__device__ __forceinline__ typename F::T jag_neutral(int q) {
  return q < 2 ? F::zero() : F::one();
}

template <class F>
struct JagLds {
  RoundLds<F, JG_NS> round;
  typename F::T pad, eq_adj;
};

// Pairs p = start, start + stride, ... of row round k: fold (FOLD: round
// k-1's rows at r into round k's state) and add W B(even), W B(even + odd)
// and W to acc. Every thread of the block calls it (it stages the round's
// cumulative pair counts in `cum`, LDS of JG_LDS_SEGS words, when they fit).
template <class F, bool FOLD>
__device__ __forceinline__ void jag_row_pairs(const JagArgs<F>& a, uint32_t k, typename F::T lam,
                                              typename F::T r, uint32_t start, uint32_t stride,
                                              typename F::T (&acc)[JG_NS], uint32_t* cum) {
  using T = typename F::T;
  const JagSeg* seg = a.segs + (size_t)k * a.nseg;
  const uint32_t P = seg[a.nseg - 1].cum_pairs;
  const bool staged = a.nseg <= JG_LDS_SEGS;
  if (staged) {
    for (uint32_t i = threadIdx.x; i < a.nseg; i += blockDim.x) cum[i] = seg[i].cum_pairs;
    __syncthreads();
  }
  const T* E = a.erow + (((size_t)1 << (a.nrv - k - 1)) - 1);  // E_{nrv-k-1}
  const T* eint = a.eint + (((size_t)1 << a.niv) - 1);         // E_niv of the batch
  // Round k - 1's state: the input for k = 1, else state[(k - 1) & 1].
  const bool from_in = !FOLD || k == 1;
  const T* src = from_in ? a.in : a.state[(k - 1) & 1];
  const size_t sstr = from_in ? a.in_stride : a.state_stride;
  T* dst = a.state[k & 1];
  for (uint32_t p = start; p < P; p += stride) {
    uint32_t lo = 0, hi = a.nseg - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if ((staged ? cum[mid] : seg[mid].cum_pairs) > p) hi = mid;
      else lo = mid + 1;
    }
    const JagSeg sg = seg[lo];
    const uint32_t j = p - (sg.cum_pairs - ((sg.count + 1) >> 1));
    T e[4], o[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const T* sq = src + (size_t)q * sstr;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const uint32_t i = 2 * j + h;
        T v = jag_neutral<F>(q);
        if (i < sg.count) {
          if constexpr (FOLD) {
            const T x = sq[sg.prev_base + 2 * i];
            const T y = 2 * i + 1 < sg.prev_count ? sq[sg.prev_base + 2 * i + 1] : jag_neutral<F>(q);
            v = F::add(x, F::mul(r, F::sub(y, x)));
            dst[(size_t)q * a.state_stride + sg.base + i] = v;
          } else {
            v = sq[sg.base + i];
          }
        }
        if (h == 0) e[q] = v;
        else o[q] = v;
      }
    }
    const T W = F::mul(E[j], eint[lo]);
    acc[0] = F::add(acc[0], F::mul(W, jag_bracket<F>(lam, e[0], e[1], e[2], e[3])));
    acc[1] = F::add(acc[1], F::mul(W, jag_bracket<F>(lam, F::add(e[0], o[0]), F::add(e[1], o[1]),
                                                         F::add(e[2], o[2]), F::add(e[3], o[3]))));
    acc[2] = F::add(acc[2], W);
  }
}

// Row round k as 1 + nblk blocks (see the file comment); block 0 warms the
// hop on a scratch copy of the transcript state while the others stream.
template <class F, bool FOLD>
__global__ __launch_bounds__(PASS_THREADS) void jagged_row_kernel(JagArgs<F> a, uint32_t k) {
  using T = typename F::T;
  __shared__ JagLds<F> L;
  __shared__ HopLds warm;
  __shared__ uint32_t cum[JG_LDS_SEGS];
  auto& sh = L.round;
  const uint32_t tid = threadIdx.x, nblk = gridDim.x - 1;
  JagCtl<F>* jc = a.ctl;
  Ctl<F>* ctl = &jc->c;
  if (blockIdx.x == 0) {
    if (tid < 64) {
      fin_load(sh.sm, ctl);
      const uint32_t* from = reinterpret_cast<const uint32_t*>(static_cast<HopLds*>(&sh.sm));
      uint32_t* to = reinterpret_cast<uint32_t*>(&warm);
      for (uint32_t i = tid; i < sizeof(HopLds) / 4; i += 64u) to[i] = from[i];
      wave_sync();
      fs_hop<HOP_PASS>((ZS_LDS HopLds*)&warm, (const ZS_LDS uint8_t*)sh.sm.msg,
                       4u * (uint32_t)sizeof(T), ctl->item_bytes, ctl->limbs, ctl->tr_kind);
    }
    if (tid == 0) {
      sh.is_last = chain_wait(&ctl->counter, nblk, &ctl->fault) ? 1u : 0u;
      L.pad = jc->pad;
      L.eq_adj = jc->eq_adj;
    }
    __syncthreads();
    if (!sh.is_last) return;
    slot_rows_sum<F, JG_NS>(a.slots, sh);
    if (tid < 64) {
      T pad = L.pad, eq_adj = L.eq_adj;
      jag_finish<F, HOP_PASS>(a, k, sh.out, sh.sm, pad, eq_adj);
      fin_store(sh.sm, ctl);
      if (tid == 0) {
        jc->pad = pad;
        jc->eq_adj = eq_adj;
      }
    }
    if (tid == 0) __hip_atomic_store(&ctl->counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  const uint32_t blk = blockIdx.x - 1;
  T acc[JG_NS] = {F::zero(), F::zero(), F::zero()};
  jag_row_pairs<F, FOLD>(a, k, ctl->lam, ctl->r, blk * PASS_THREADS + tid, nblk * PASS_THREADS,
                         acc, cum);
  block_reduce<F, JG_NS, PASS_THREADS>(acc, sh.lds, sh.out);
  slot_row_add<F, JG_NS>(a.slots, blk, sh.out);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (tid == 0) __hip_atomic_fetch_add(&ctl->counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One block: row rounds kfirst..nrv-1 (those whose pairs are too few to
// fill the device: a launch and a device-wide arrival each would cost more
// than the pairs), then the last row round's state folded at its challenge
// into one element a segment, each batch round (sum its pairs, finish, fold)
// and the final fold.
template <class F>
__global__ __launch_bounds__(PASS_THREADS) void jagged_batch_kernel(JagArgs<F> a,
                                                                    uint32_t kfirst) {
  using T = typename F::T;
  __shared__ JagLds<F> L;
  __shared__ uint32_t cum[JG_LDS_SEGS];
  auto& sh = L.round;
  const uint32_t tid = threadIdx.x;
  JagCtl<F>* jc = a.ctl;
  Ctl<F>* ctl = &jc->c;
  if (tid < 64) fin_load(sh.sm, ctl);
  if (tid == 0) {
    L.pad = jc->pad;
    L.eq_adj = jc->eq_adj;
  }
  __syncthreads();
  const T lam = ctl->lam;
  for (uint32_t k = kfirst; k < a.nrv; ++k) {
    T acc[JG_NS] = {F::zero(), F::zero(), F::zero()};
    if (k == 0)
      jag_row_pairs<F, false>(a, k, lam, sh.sm.r, tid, PASS_THREADS, acc, cum);
    else
      jag_row_pairs<F, true>(a, k, lam, sh.sm.r, tid, PASS_THREADS, acc, cum);
    block_reduce<F, JG_NS, PASS_THREADS>(acc, sh.lds, sh.out);
    if (tid < 64) {
      T pad = L.pad, eq_adj = L.eq_adj;
      jag_finish<F, HOP_INLINE>(a, k, sh.out, sh.sm, pad, eq_adj);
      if (tid == 0) {
        L.pad = pad;
        L.eq_adj = eq_adj;
      }
    }
    __syncthreads();
  }
  // Boundary fold at r_{nrv-1}: segment s's rows 0, 1 of the last row round.
  {
    const uint32_t k = a.nrv - 1;
    const JagSeg* seg = a.segs + (size_t)k * a.nseg;
    const T* src = k == 0 ? a.in : a.state[k & 1];
    const size_t sstr = k == 0 ? a.in_stride : a.state_stride;
    const T r = sh.sm.r;
    for (uint32_t s = tid; s < a.nseg; s += PASS_THREADS) {
      const JagSeg sg = seg[s];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const T* sq = src + (size_t)q * sstr;
        const T x = sq[sg.base];
        const T y = sg.count >= 2 ? sq[sg.base + 1] : jag_neutral<F>(q);
        a.dense[0][(size_t)q * a.dense_stride + s] = F::add(x, F::mul(r, F::sub(y, x)));
      }
    }
  }
  __syncthreads();
  for (uint32_t m = 0; m < a.niv; ++m) {
    const uint32_t len = a.nseg >> m, half = len >> 1;
    const T* D = a.dense[m & 1];
    const T* E = a.eint + (((size_t)1 << (a.niv - m - 1)) - 1);  // E_{niv-m-1}
    T acc[JG_NS] = {F::zero(), F::zero(), F::zero()};
    for (uint32_t b = tid; b < half; b += PASS_THREADS) {
      T e[4], o[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        e[q] = D[(size_t)q * a.dense_stride + 2 * b];
        o[q] = D[(size_t)q * a.dense_stride + 2 * b + 1];
      }
      const T W = E[b];
      acc[0] = F::add(acc[0], F::mul(W, jag_bracket<F>(lam, e[0], e[1], e[2], e[3])));
      acc[1] = F::add(acc[1], F::mul(W, jag_bracket<F>(lam, F::add(e[0], o[0]), F::add(e[1], o[1]),
                                                           F::add(e[2], o[2]), F::add(e[3], o[3]))));
      acc[2] = F::add(acc[2], W);
    }
    block_reduce<F, JG_NS, PASS_THREADS>(acc, sh.lds, sh.out);
    if (tid < 64) {
      T pad = L.pad, eq_adj = L.eq_adj;
      jag_finish<F, HOP_INLINE>(a, a.nrv + m, sh.out, sh.sm, pad, eq_adj);
      if (tid == 0) {
        L.pad = pad;
        L.eq_adj = eq_adj;
      }
    }
    __syncthreads();
    const T r = sh.sm.r;
    T* Dn = a.dense[(m + 1) & 1];
    for (uint32_t b = tid; b < half; b += PASS_THREADS) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const T x = D[(size_t)q * a.dense_stride + 2 * b];
        const T y = D[(size_t)q * a.dense_stride + 2 * b + 1];
        Dn[(size_t)q * a.dense_stride + b] = F::add(x, F::mul(r, F::sub(y, x)));
      }
    }
    __syncthreads();
  }
  if (tid < 4) a.openings[tid] = a.dense[a.niv & 1][(size_t)tid * a.dense_stride];
  if (tid < 64) fin_store(sh.sm, ctl);
  if (tid == 0) {
    jc->pad = L.pad;
    jc->eq_adj = L.eq_adj;
  }
}

template <class F>
static int prove_jagged_typed(zs_ctx* ctx, const zs_jagged_args& a, zs_transcript* t,
                              zs_jagged_proof* out) {
  using T = typename F::T;
  const uint32_t nseg = a.num_segments, nrv = a.row_variables;
  uint32_t niv = 0;
  while ((1u << niv) < nseg) ++niv;
  const uint32_t R = nrv + niv;
  // Row schedules: counts_k = ceil(counts_{k-1} / 2), pairs_k = ceil(counts_k / 2).
  std::vector<JagSeg> segs((size_t)nrv * nseg);
  std::vector<uint32_t> cnt(a.row_counts, a.row_counts + nseg), prev(nseg, 0), pbase(nseg, 0);
  uint64_t width = 0;
  for (uint32_t s = 0; s < nseg; ++s) width += cnt[s];
  size_t state_len = 1;
  for (uint32_t k = 0; k < nrv; ++k) {
    if (k > 0)
      for (uint32_t s = 0; s < nseg; ++s) cnt[s] = (prev[s] + 1) / 2;
    uint32_t cum = 0, base = 0;
    for (uint32_t s = 0; s < nseg; ++s) {
      JagSeg& g = segs[(size_t)k * nseg + s];
      cum += (cnt[s] + 1) / 2;
      g.cum_pairs = cum;
      g.base = base;
      g.count = cnt[s];
      g.prev_base = k ? pbase[s] : 0;
      g.prev_count = k ? prev[s] : 0;
      g.pad_[0] = g.pad_[1] = g.pad_[2] = 0;
      base += cnt[s];
    }
    if (k > 0 && base > state_len) state_len = base;
    for (uint32_t s = 0; s < nseg; ++s) {
      pbase[s] = segs[(size_t)k * nseg + s].base;
      prev[s] = cnt[s];
    }
  }
  // Round constants.
  std::vector<JagRound<F>> rcs(R);
  const T one = F::one(), two = F::from_int(2), half = F::inv(two);
  for (uint32_t rnd = 0; rnd < R; ++rnd) {
    const T z = load_host_el<F>(a.eval_point, R - 1 - rnd);
    const T omz = F::sub(one, z);
    const T den = F::sub(one, F::mul(two, z));
    if (F::eq(F::canon(den), F::zero()))
      return fail(ZS_EINVAL, "eval_point coordinate 1/2: the round's eq factor has no root");
    const std::vector<T> xs = {F::zero(), one, half, F::mul(omz, F::inv(den))};
    FMat<F> vinv;
    if (!inv_vandermonde<F>(xs, vinv))
      return fail(ZS_EINVAL, "eval_point coordinate makes the interpolation nodes collide");
    rcs[rnd].z = z;
    rcs[rnd].omz = omz;
    for (int j = 0; j < 4; ++j)
      for (int i = 0; i < 3; ++i) rcs[rnd].m[j][i] = vinv[j][i];
  }
  // One pool block: io (uploaded, outputs downloaded) then device-only parts.
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off = align256(off + bytes); return o; };
  const size_t o_ctl = take(sizeof(JagCtl<F>));
  const size_t o_msgs = take((size_t)R * 4 * sizeof(T));
  const size_t o_point = take((size_t)R * sizeof(T));
  const size_t o_open = take(4 * sizeof(T));
  const size_t out_bytes = off;
  const size_t o_slots = take((size_t)NSLOT * SLOTW * sizeof(uint64_t));
  const size_t o_segs = take(segs.size() * sizeof(JagSeg));
  const size_t o_rc = take((size_t)R * sizeof(JagRound<F>));
  const size_t o_z = take((size_t)R * sizeof(T));
  const size_t io_bytes = off;
  const size_t o_erow = take((((size_t)1 << nrv) - 1) * sizeof(T));
  const size_t o_eint = take((((size_t)2 << niv) - 1) * sizeof(T));
  const size_t o_state = take((size_t)2 * 4 * state_len * sizeof(T));
  const size_t o_dense = take((size_t)2 * 4 * nseg * sizeof(T));
  const bool stage = !a.planes_on_device;
  const size_t o_in = stage ? take((size_t)4 * width * sizeof(T)) : 0;
  void* block = nullptr;
  void* host = nullptr;
  int rc;
  if ((rc = ctx->pool->alloc(off, &block))) return rc;
  if ((rc = ctx->host_pool->alloc(io_bytes, &host))) {
    ctx->pool->release(block);
    return rc;
  }
  struct Release {
    zs_ctx* c;
    void* b;
    void* h;
    ~Release() {
      c->pool->release(b);
      c->host_pool->release(h);
    }
  } release{ctx, block, host};
  uint8_t* b = (uint8_t*)block;
  uint8_t* h = (uint8_t*)host;
  memset(h, 0, io_bytes);
  JagCtl<F>& jc = *(JagCtl<F>*)(h + o_ctl);
  Ctl<F>& c = jc.c;
  memcpy(&c.tr, &t->sha, sizeof(TrState));
  c.tr_kind = t->kind;
  c.item_bytes = t->item_bytes;
  c.limbs = (uint32_t)(sizeof(T) / t->item_bytes);
  c.claim = load_host_el<F>(a.claim, 0);
  c.r = F::zero();
  c.eq_prev = one;
  c.one = one;
  c.wire = WIRE_VALUES;
  c.mlen = 4;
  c.lam = load_host_el<F>(a.lambda, 0);
  jc.pad = one;
  jc.eq_adj = one;
  jc.inv8 = F::inv(F::from_int(8));
  jc.four = F::from_int(4);
  memcpy(h + o_segs, segs.data(), segs.size() * sizeof(JagSeg));
  memcpy(h + o_rc, rcs.data(), rcs.size() * sizeof(JagRound<F>));
  T* zh = (T*)(h + o_z);
  for (uint32_t i = 0; i < R; ++i) zh[i] = load_host_el<F>(a.eval_point, i);  // batch || row
  JagArgs<F> g{};
  g.in = stage ? (const T*)(b + o_in) : (const T*)a.planes;
  g.in_stride = width;
  T* st = (T*)(b + o_state);
  g.state[0] = st;
  g.state[1] = st + (size_t)4 * state_len;
  g.state_stride = state_len;
  T* dn = (T*)(b + o_dense);
  g.dense[0] = dn;
  g.dense[1] = dn + (size_t)4 * nseg;
  g.dense_stride = nseg;
  g.segs = (const JagSeg*)(b + o_segs);
  g.rc = (const JagRound<F>*)(b + o_rc);
  g.erow = (const T*)(b + o_erow);
  g.eint = (const T*)(b + o_eint);
  g.zint = (const T*)(b + o_z);
  g.zrow = (const T*)(b + o_z) + niv;
  g.ctl = (JagCtl<F>*)(b + o_ctl);
  g.slots = (uint64_t*)(b + o_slots);
  g.msgs = (T*)(b + o_msgs);
  g.point = b + o_point;
  g.openings = (T*)(b + o_open);
  g.nseg = nseg;
  g.nrv = nrv;
  g.niv = niv;
  hipStream_t s = ctx->stream;
  if (stage)
    ZS_HIP(hipMemcpyAsync((void*)g.in, a.planes, (size_t)4 * width * sizeof(T),
                          hipMemcpyHostToDevice, s));
  ZS_HIP(hipMemcpyAsync(b, h, io_bytes, hipMemcpyHostToDevice, s));
  auto eq_grid = [](size_t n) {
    size_t gsz = (n + PASS_THREADS - 1) / PASS_THREADS;
    return (uint32_t)(gsz < 1 ? 1 : gsz > 1024 ? 1024 : gsz);
  };
  // Row family E_0..E_{nrv-1}, batch family E_0..E_niv.
  jagged_eq_kernel<F><<<eq_grid(((size_t)1 << nrv) - 1), PASS_THREADS, 0, s>>>(
      g.zrow, nrv - 1, (T*)g.erow);
  ZS_HIP(hipGetLastError());
  jagged_eq_kernel<F><<<eq_grid(((size_t)2 << niv) - 1), PASS_THREADS, 0, s>>>(g.zint, niv,
                                                                                  (T*)g.eint);
  ZS_HIP(hipGetLastError());
  // Kernel timing (zs_ctx_profile): algorithmic bytes of the four planes.
  auto begin = [&]() -> hipEvent_t {
    if (!ctx->profile) return nullptr;
    hipEvent_t e = ctx->event();
    (void)hipEventRecord(e, s);
    return e;
  };
  auto mark = [&](hipEvent_t e, uint32_t round, uint32_t kind, uint32_t grid, uint64_t bytes) {
    if (!e) return;
    hipEvent_t stop = ctx->event();
    (void)hipEventRecord(stop, s);
    ctx->records.push_back(KernelRecord{round, kind, grid, bytes, e, stop});
  };
  uint64_t live_prev = width;
  const uint32_t maxg = ctx->tune.max_grid < (uint32_t)MAXG ? ctx->tune.max_grid : (uint32_t)MAXG;
  uint32_t kfirst = 0;
  for (uint32_t k = 0; k < nrv; ++k) {
    const JagSeg& last = segs[(size_t)k * nseg + nseg - 1];
    const uint32_t P = last.cum_pairs;
    const uint64_t live = (uint64_t)last.base + last.count;
    if (P <= JG_TAIL_PAIRS) break;  // this round and the rest: one block
    kfirst = k + 1;
    size_t grid = (P + PASS_THREADS - 1) / PASS_THREADS;
    if (grid > maxg) grid = maxg;
    if (grid < 1) grid = 1;
    hipEvent_t ev = begin();
    if (k == 0)
      jagged_row_kernel<F, false><<<(uint32_t)grid + 1, PASS_THREADS, 0, s>>>(g, k);
    else
      jagged_row_kernel<F, true><<<(uint32_t)grid + 1, PASS_THREADS, 0, s>>>(g, k);
    ZS_HIP(hipGetLastError());
    // Round 0 reads the live rows; round k reads round k-1's and writes its own.
    mark(ev, k, KREC_JAGGED, (uint32_t)grid,
         (uint64_t)4 * sizeof(T) * (k == 0 ? live : live_prev + live));
    live_prev = live;
  }
  hipEvent_t evb = begin();
  jagged_batch_kernel<F><<<1, PASS_THREADS, 0, s>>>(g, kfirst);
  ZS_HIP(hipGetLastError());
  // Row rounds read the previous round's rows and write their own (round 0
  // only reads); the batch phase reads the last row state once.
  uint64_t tail_rows = 2 * (uint64_t)nseg;
  for (uint32_t k = kfirst; k < nrv; ++k) {
    const JagSeg& last = segs[(size_t)k * nseg + nseg - 1];
    const uint64_t live = (uint64_t)last.base + last.count;
    tail_rows += k == 0 ? live : live_prev + live;
    live_prev = live;
  }
  tail_rows += live_prev;
  const uint64_t tail_bytes = (uint64_t)4 * sizeof(T) * tail_rows;
  mark(evb, kfirst, KREC_JAGGED_BATCH, 1, tail_bytes);
  ZS_HIP(hipMemcpyAsync(h, b, out_bytes, hipMemcpyDeviceToHost, s));
  ZS_HIP(hipStreamSynchronize(s));
  const JagCtl<F>& jo = *(const JagCtl<F>*)(h + o_ctl);
  if (jo.c.fault) return fail(ZS_EHIP, "jagged rounds: a device wait reached its limit");
  if (out->round_polys) memcpy(out->round_polys, h + o_msgs, (size_t)R * 4 * sizeof(T));
  if (out->challenges) memcpy(out->challenges, h + o_point, (size_t)R * sizeof(T));
  if (out->openings) memcpy(out->openings, h + o_open, 4 * sizeof(T));
  if (out->value) memcpy(out->value, &jo.c.claim, sizeof(T));
  memcpy(&t->sha, &jo.c.tr, sizeof(TrState));
  return ZS_OK;
}

// One transition of the jagged circuit on the device (reference
// circuit.py:296-470): output slot q of segment s (binary search over the
// out counts' prefix sums) is row j of the segment; while j < ceil(rc/2) it
// folds the segment's rows 2j and 2j + 1 (the odd one the neutral fraction
// past the count), later slots of the out count fold two neutral rows, and
// slots past the live out rows are zero. Planes n0, n1, d0, d1, row stride
// = the width.
struct JagTrSeg {
  uint32_t cum_out;  // out rows of segments 0..s (inclusive)
  uint32_t out;      // out rows of segment s
  uint32_t base;     // first row of segment s in the input
  uint32_t rows;     // input rows of segment s
};

template <class F>
__global__ __launch_bounds__(PASS_THREADS) void jagged_transition_kernel(
    const JagTrSeg* seg, uint32_t nseg, const typename F::T* in, size_t in_stride,
    typename F::T* out, size_t out_stride, uint32_t out_width) {
  using T = typename F::T;
  const uint32_t live = seg[nseg - 1].cum_out;
  for (uint32_t q = blockIdx.x * PASS_THREADS + threadIdx.x; q < out_width;
       q += gridDim.x * PASS_THREADS) {
    T v[4][2];  // plane, even / odd source
    if (q >= live) {
#pragma unroll
      for (int p = 0; p < 4; ++p) out[(size_t)p * out_stride + q] = F::zero();
      continue;
    }
    uint32_t lo = 0, hi = nseg - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (seg[mid].cum_out > q) hi = mid;
      else lo = mid + 1;
    }
    const JagTrSeg sg = seg[lo];
    const uint32_t j = q - (sg.cum_out - sg.out);
    const bool folds = j < (sg.rows + 1) / 2;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t r = 2 * j + h;
      const bool real = folds && r < sg.rows;
#pragma unroll
      for (int p = 0; p < 4; ++p)
        v[p][h] = real ? in[(size_t)p * in_stride + sg.base + r] : jag_neutral<F>(p);
    }
    // n0' = n0 d1 + n1 d0 and d0' = d0 d1 at the even row; n1', d1' at the odd.
    out[0 * out_stride + q] = F::add(F::mul(v[0][0], v[3][0]), F::mul(v[1][0], v[2][0]));
    out[1 * out_stride + q] = F::add(F::mul(v[0][1], v[3][1]), F::mul(v[1][1], v[2][1]));
    out[2 * out_stride + q] = F::mul(v[2][0], v[3][0]);
    out[3 * out_stride + q] = F::mul(v[2][1], v[3][1]);
  }
}

template <class F>
static int jagged_transition_typed(zs_ctx* ctx, uint32_t nseg, const uint32_t* rows,
                                   const uint32_t* out_rows, const void* planes, size_t width,
                                   void* next, size_t out_width) {
  using T = typename F::T;
  std::vector<JagTrSeg> segs(nseg);
  uint32_t cum = 0, base = 0;
  for (uint32_t s = 0; s < nseg; ++s) {
    cum += out_rows[s];
    segs[s] = JagTrSeg{cum, out_rows[s], base, rows[s]};
    base += rows[s];
  }
  void* d = nullptr;
  int rc;
  if ((rc = ctx->pool->alloc(nseg * sizeof(JagTrSeg), &d))) return rc;
  hipError_t e = hipMemcpyAsync(d, segs.data(), nseg * sizeof(JagTrSeg), hipMemcpyHostToDevice,
                                ctx->stream);
  if (e == hipSuccess) {
    size_t g = (out_width + PASS_THREADS - 1) / PASS_THREADS;
    g = g < 1 ? 1 : g > 4096 ? 4096 : g;
    jagged_transition_kernel<F><<<(uint32_t)g, PASS_THREADS, 0, ctx->stream>>>(
        (const JagTrSeg*)d, nseg, (const T*)planes, width, (T*)next, out_width,
        (uint32_t)out_width);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx->pool->release(d);
  if (e != hipSuccess) return hip_fail(e, "jagged transition");
  return ZS_OK;
}

int jagged_transition(zs_ctx* ctx, uint32_t field, uint32_t nseg, const uint32_t* rows,
                      const uint32_t* out_rows, const void* planes, size_t width, void* next,
                      size_t out_width) {
  switch (field) {
    case ZS_F_BABYBEAR:
      return jagged_transition_typed<BB>(ctx, nseg, rows, out_rows, planes, width, next, out_width);
    case ZS_F_BABYBEAR4:
      return jagged_transition_typed<BB4>(ctx, nseg, rows, out_rows, planes, width, next, out_width);
    case ZS_F_GOLDILOCKS:
      return jagged_transition_typed<GL>(ctx, nseg, rows, out_rows, planes, width, next, out_width);
    default: return fail(ZS_EINVAL, "unknown field");
  }
}

int prove_jagged_layer(zs_ctx* ctx, const zs_jagged_args& a, zs_transcript* t,
                       zs_jagged_proof* out) {
  switch (a.field) {
    case ZS_F_BABYBEAR: return prove_jagged_typed<BB>(ctx, a, t, out);
    case ZS_F_BABYBEAR4: return prove_jagged_typed<BB4>(ctx, a, t, out);
    case ZS_F_GOLDILOCKS: return prove_jagged_typed<GL>(ctx, a, t, out);
    default: return fail(ZS_EINVAL, "unknown field");
  }
}

}  // namespace zs
