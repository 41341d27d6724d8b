// Dense sumcheck round kernels for gfx950 (CDNA4, wave64).
//
// One round of the reference's StandardRound / EqPolyRound
// (zorch/sumcheck/prover.py:157-174, zorch/sumcheck/eq/eq_poly.py:114-221) is
// one streaming pass over HBM that
//   1. folds the previous round's tables at the previous challenge,
//      T'[x] = T[x] + r (T[x + L] - T[x])         (domain.py:226-233, MSB bind),
//   2. evaluates this round's polynomial on the natural nodes t = 0..K,
//      t(t) = sum_j w(j) prod_k (lo_k[j] + t (hi_k[j] - lo_k[j]))
//                                               (domain.py:131-137,158-205),
//      with w the factored eq weight eq_l[j_hi] eq_r[j_lo] in eq mode
//      (eq_poly.py:63-97,144-187), never materialised.
// Each block adds its node sums, word by word, into one of NSLOT 64-bit slot
// rows with device-scope atomics (field words are < 2^32, so a slot never
// overflows); the last block to finish (a ticket counter) sums the slot rows,
// reduces the words mod p and finishes the round on one wavefront, one node
// per lane: it extends t to node K+1 in eq mode (a degree-K polynomial is
// fixed by K+1 values, so t(K+1) = sum_i (-1)^(K-i) C(K+1, i) t(i) exactly),
// forms the message s(t) = l(t) t(t) (eq_poly.py:100-111), runs the
// Fiat-Shamir hop of the proof's transcript (SHA-256: hop.h, BLAKE3:
// blake3.h), binds the challenge (round.py:41-49) and, where the
// claim is read, reduces it by Lagrange interpolation (reduce.py:27-52,
// poly/univariate.py:36-65). The challenge stays in device memory, so a whole
// stage is a chain of launches on one stream with no host round trip.
//
// Layout: tables are table-major (table q at q*stride), element-contiguous;
// BabyBear^4 elements are 16 B and move as one dwordx4. Pair j of a round with
// table length L reads T[j] and T[j + L/2]; a folding pass reads the four
// quarter-strided source entries j + {0, L/2, L, 3L/2} and writes the two
// folded entries j and j + L/2 — only positions thread j itself read, so the
// fold runs in place after the first round. Consecutive lanes touch
// consecutive addresses in all four streams.
//
// Eq weights: with a power-of-two grid whose thread count is a multiple of
// |eq_r|, every pair a thread visits has the same low index bits, so eq_r is
// a per-thread constant applied once after the loop and each pair multiplies
// by eq_l only. When the tables are already in the challenge field the weight
// is folded into the first factor (w lo_0, w hi_0) before the node products.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "blake3.h"
#include "field.h"
#include "hop.h"
#include "sha256.h"

namespace zs {

constexpr int MAXPTS = 6;     // degree <= 4 product + eq factor
constexpr int MAXK = 4;
constexpr int PASS_THREADS = 256;
constexpr int TAIL_THREADS = 512;
constexpr int MAXG = 4096;    // max blocks of one pass
constexpr int NSLOT = 32;     // slot rows of the cross-block node sums
constexpr int SLOTW = 32;     // 64-bit words per slot row (>= (K+1) x words/element)

// Transcript kinds (the C ABI's ZS_TR_*) and their streaming states.
constexpr uint32_t TR_SHA256 = 1, TR_BLAKE3 = 2;
union TrState {
  Sha256State sha;
  Blake3State b3;
};

// Device control block of one proof (lives in device memory).
template <class FC>
struct Ctl {
  TrState tr;            // transcript stream
  uint32_t tr_kind;      // TR_SHA256 or TR_BLAKE3
  uint32_t item_bytes;   // transcript word size
  uint32_t limbs;        // transcript words per challenge
  uint32_t counter;      // last-block ticket of the current pass; chained arrivals
  uint32_t done;         // rounds published by a chained launch
  uint32_t fault;        // a chained wait reached its limit
  uint32_t pad_;
  typename FC::T claim;
  typename FC::T r;        // current challenge, canonical
  typename FC::T eq_prev;  // eq mass of the bound prefix (eq mode)
  typename FC::T one;
  typename FC::T lag_inv[MAXPTS];  // 1 / prod_{j != t} (t - j) over {0..npts-1}
  typename FC::T nodes[MAXPTS];    // t = 0..npts-1 as field elements
  typename FC::T ext_coef[MAXPTS]; // (-1)^(K-i) C(K+1, i): t(K+1) from t(0..K)
  // Round-message wire form (the C ABI's ZS_WIRE_*), with its maps built on
  // the host: message element j is sum_i wmat[j][i] s(i) over the natural
  // values s(0..npts-1), and the reduced claim is sum_k r^k c_k with
  // c_k = sum_j cmat[k][j] msg_j + cclaim[k] claim. WIRE_VALUES (the natural
  // values themselves) skips both maps and reduces by Lagrange.
  uint32_t wire;
  uint32_t mlen;         // message elements (<= npts)
  uint32_t ncoef;        // coefficients of the reduced polynomial
  uint32_t claim_every;  // the reduction reads the running claim: reduce every round
  typename FC::T wmat[MAXPTS][MAXPTS];
  typename FC::T cmat[MAXPTS][MAXPTS];
  typename FC::T cclaim[MAXPTS];
  typename FC::T lam;  // LogUp stages: the layer's batching challenge
};
constexpr uint32_t WIRE_VALUES = 0, WIRE_DOMAIN = 1, WIRE_COEFFS = 2, WIRE_COMPRESSED = 3;

template <class FT, class FC>
struct StageArgs {
  const typename FT::T* in;     // input tables (k x 2^n), table field
  size_t in_stride;
  typename FC::T* work;         // folded tables (k x 2^(n-1)), challenge field
  size_t work_stride;
  uint32_t n;
  uint32_t k;
  Ctl<FC>* ctl;
  uint64_t* slots;              // NSLOT x SLOTW, all zero between rounds
  typename FC::T* msgs;         // n x npts
  uint8_t* point;               // n x sizeof(FC::T) raw challenge bytes
  typename FC::T* final_out;    // k (+1) values, may be null
  const typename FC::T* eq_point;  // n coordinates (eq mode)
  const typename FC::T* eq_ab;     // n x (2 w_i - 1, 1 - w_i): l_i(t) = eq_prev (B + t A)
  const typename FC::T* eq_l_fam;  // suffix family of eq_point[0 : n/2]
  const typename FC::T* eq_r_fam;  // suffix family of eq_point[n/2 : n]
  uint64_t* trace;                 // phase timestamps (null: off)
  typename FC::T* sq_eq;           // sqrt-space engine: two eq(r[<i], .) buffers of 2^sq_half
  uint32_t sq_half;                // sqrt-space engine: refold rounds (0: none)
};

// Phase trace: trace[0] counts events; event i is (round << 8 | slot,
// wall-clock ticks) at trace[1 + 2i]. One lane records; off when null.
__device__ __forceinline__ void trace_mark(uint64_t* trace, uint32_t round, uint32_t slot) {
  if (trace) {
    const uint64_t i = __hip_atomic_fetch_add(trace, (uint64_t)1, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
    if (i < 2047) {
      trace[1 + 2 * i] = ((uint64_t)round << 8) | slot;
      trace[2 + 2 * i] = wall_clock64();
    }
  }
}

// --------------------------------------------------- challenge multiplier
// Multiplication by the round challenge r, with the X^4 = 11 wrap of r
// precomputed once per thread for BabyBear^4.
template <class FC>
struct ChalMul {
  typename FC::T r;
  __device__ __forceinline__ void init(typename FC::T v) { r = v; }
  __device__ __forceinline__ typename FC::T mul(typename FC::T x) const { return FC::mul(x, r); }
};
template <>
struct ChalMul<BB4> {
  bb4 r;
  uint32_t w1, w2, w3;
  __device__ __forceinline__ void init(bb4 v) {
    r = v;
    w1 = bb_mul11(v.c[1]);
    w2 = bb_mul11(v.c[2]);
    w3 = bb_mul11(v.c[3]);
  }
  __device__ __forceinline__ bb4 mul(bb4 x) const { return bb4_mul_w(x, r, w1, w2, w3); }
};

template <class FC, class FS>
struct FoldMul {
  static __device__ __forceinline__ typename FC::T go(const ChalMul<FC>& cm, typename FS::T x) {
    return cm.mul(x);
  }
};
template <>
struct FoldMul<BB4, BB> {
  static __device__ __forceinline__ bb4 go(const ChalMul<BB4>& cm, uint32_t x) {
    return bb4_scale(cm.r, x);
  }
};

// ------------------------------------------------------------ reductions
template <class F>
__device__ __forceinline__ typename F::T shfl_xor_el(typename F::T v, int m);
template <>
__device__ __forceinline__ uint32_t shfl_xor_el<BB>(uint32_t v, int m) {
  return __shfl_xor(v, m, 64);
}
template <>
__device__ __forceinline__ bb4 shfl_xor_el<BB4>(bb4 v, int m) {
  return bb4_make(__shfl_xor(v.c[0], m, 64), __shfl_xor(v.c[1], m, 64),
                  __shfl_xor(v.c[2], m, 64), __shfl_xor(v.c[3], m, 64));
}
template <>
__device__ __forceinline__ uint64_t shfl_xor_el<GL>(uint64_t v, int m) {
  uint32_t lo = __shfl_xor((uint32_t)v, m, 64), hi = __shfl_xor((uint32_t)(v >> 32), m, 64);
  return ((uint64_t)hi << 32) | lo;
}

// Block-wide field sum of N accumulators; the result lands in out[] (LDS)
// and is valid for every thread after the trailing barrier.
template <class F, int N, int THREADS>
__device__ __forceinline__ void block_reduce(typename F::T (&acc)[N], typename F::T* lds, typename F::T* out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < N; ++t) {
    typename F::T v = acc[t];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = F::add(v, shfl_xor_el<F>(v, m));
    acc[t] = v;
  }
  if (lane == 0) {
#pragma unroll
    for (int t = 0; t < N; ++t) lds[wave * N + t] = acc[t];
  }
  __syncthreads();
  if (threadIdx.x < N) {
    typename F::T s = F::zero();
    for (int w = 0; w < THREADS / 64; ++w) s = F::add(s, lds[w * N + threadIdx.x]);
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// Sum of the (lane & 7) groups of eight lanes: every lane of a group gets the
// field sum of the group's values.
template <class F>
__device__ __forceinline__ typename F::T lanes8_sum(typename F::T v) {
#pragma unroll
  for (int m = 1; m < 8; m <<= 1) v = F::add(v, shfl_xor_el<F>(v, m));
  return v;
}

// Cross-block node sums as 32-bit words added into 64-bit slots: an element
// is W words, each below 2^32, and a slot total is reduced back to the field.
// Sums of BabyBear Montgomery words are Montgomery words of the sum.
__device__ __forceinline__ uint32_t bb_from_u64(uint64_t x) {  // x < p 2^32
  return bb_mul(bb_redc(x), BB_R2);                             // (x / 2^32) 2^64 / 2^32
}
template <class F>
struct SlotOps;
template <>
struct SlotOps<BB> {
  static constexpr int W = 1;
  static __device__ __forceinline__ uint32_t word(uint32_t v, int) { return v; }
  static __device__ __forceinline__ uint32_t join(const uint64_t* s) { return bb_from_u64(s[0]); }
};
template <>
struct SlotOps<BB4> {
  static constexpr int W = 4;
  static __device__ __forceinline__ uint32_t word(bb4 v, int c) { return v.c[c]; }
  static __device__ __forceinline__ bb4 join(const uint64_t* s) {
    return bb4_make(bb_from_u64(s[0]), bb_from_u64(s[1]), bb_from_u64(s[2]), bb_from_u64(s[3]));
  }
};
template <>
struct SlotOps<GL> {
  static constexpr int W = 2;  // low and high halves
  static __device__ __forceinline__ uint32_t word(uint64_t v, int c) {
    return c ? (uint32_t)(v >> 32) : (uint32_t)v;
  }
  static __device__ __forceinline__ uint64_t join(const uint64_t* s) {  // s[i] < 2^44 < p
    return gl_add(s[0], gl_mul(s[1], (uint64_t)1 << 32));
  }
};

// ------------------------------------------------------------- eq weights
// Round index i = round + 1 (reference numbering): both half tables while
// i < n/2, then the right family alone, then weight one (eq_poly.py:144-187).
__device__ __host__ inline size_t fam_offset(uint32_t member) {
  return ((size_t)2 << member) - 2;  // member m has 2^(m+1) entries
}

template <class FC>
__device__ inline void select_eq(uint32_t round, uint32_t n, const typename FC::T* lfam,
                                 const typename FC::T* rfam, const typename FC::T* one,
                                 const typename FC::T*& wl, const typename FC::T*& wr,
                                 uint32_t& nxr_log) {
  const uint32_t i = round + 1, rest = n - i, l_half = n / 2;
  if (rest == 0) {
    wl = nullptr; wr = one; nxr_log = 0;
  } else if (i < l_half) {
    wl = lfam + fam_offset((l_half - i) - 1);
    wr = rfam + fam_offset((n - l_half) - 1);
    nxr_log = n - l_half;
  } else {
    wl = nullptr; wr = rfam + fam_offset(rest - 1); nxr_log = rest;
  }
}

// ------------------------------------------------------------ pass body
// Accumulate this thread's pairs j = start, start+stride, ... < L/2 into
// acc[t] = t(t), t = 0..K. FOLD: source tables have length 2L and are folded
// at r into dst first; else the source tables are this round's tables.
// PassUnroll pairs are loaded before any is folded or stored, so that a
// thread keeps at least 64 bytes of loads in flight whatever the element size.
// (Hoisting the loads is safe in place: the positions a thread stores, j and
// j + L/2, are never positions it loads for another pair.)
template <class FS, int K, bool FOLD>
struct PassUnroll {
  static constexpr int bytes = K * (FOLD ? 4 : 2) * (int)sizeof(typename FS::T);
  static constexpr int value = bytes >= 64 ? 1 : bytes >= 32 ? 2 : 4;
};

// Sum of w * x over a thread's pairs (w a challenge-field weight, x a value
// of the pass field). BabyBear^4 weights on BabyBear values keep four 64-bit
// limb sums of raw products (each < p^2, folded below p 2^32 after every
// addition, so nothing overflows) and reduce once at the end: one multiply-add
// per limb instead of a Montgomery product and a modular addition.
template <class FC, class FV>
struct WeightedSum {
  typename FC::T s;
  __device__ __forceinline__ void init() { s = FC::zero(); }
  __device__ __forceinline__ void add(typename FC::T w, typename FV::T x, bool = true) {
    s = FC::add(s, MixMul<FC, FV>::go(w, x));
  }
  __device__ __forceinline__ typename FC::T get() const { return s; }
};
// Two additions fit between folds: from below p 2^32, adding two products
// (each < p^2) stays below p (2^32 + 2p) < 2^64, and the fold then brings the
// high word below p again. `add(w, x, false)` skips the fold; the caller
// alternates.
template <>
struct WeightedSum<BB4, BB> {
  uint64_t s[4];
  __device__ __forceinline__ void init() { s[0] = s[1] = s[2] = s[3] = 0; }
  __device__ __forceinline__ void add(bb4 w, uint32_t x, bool fold = true) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint64_t v = s[c] + (uint64_t)w.c[c] * x;
      s[c] = fold ? bb_pre(v) : v;
    }
  }
  __device__ __forceinline__ bb4 get() const {
    return bb4_make(bb_redc(bb_pre(s[0])), bb_redc(bb_pre(s[1])), bb_redc(bb_pre(s[2])),
                    bb_redc(bb_pre(s[3])));
  }
};

// Source bytes a thread loads per step of a double-buffered streaming loop.
constexpr int PF_STEP_BYTES = 64;
template <class FS, class FC, int K, bool EQ, bool FOLD>
__device__ __forceinline__ void pass_accumulate(
    const typename FS::T* src, size_t src_stride, typename FC::T* dst, size_t dst_stride,
    uint32_t L, const ChalMul<FC>& cm, const typename FC::T* wl, const typename FC::T* wr,
    uint32_t nxr_log, size_t start, size_t stride, typename FC::T (&acc)[K + 1]) {
  using FV = typename std::conditional<FOLD, FC, FS>::type;
  using V = typename FV::T;
  using T = typename FC::T;
  using S = typename FS::T;
  // The first folding round of two base-field tables under a BabyBear^4
  // challenge r: each node value is a + r b with a, b in the base field
  // (a = lo + t (hi - lo) over the unfolded entries, b the same over their
  // differences), so the product of the two factors is c0 + r c1 + r^2 c2
  // with base-field c0 = a0 a1, c1 = a0 b1 + b0 a1, c2 = b0 b1. Summing the
  // c_i per node (weighted by eq, lazily) and applying r, r^2 once per thread
  // replaces five extension products a pair with four base ones.
  constexpr bool SPLIT = FOLD && K == 2 && std::is_same<FS, BB>::value &&
                         std::is_same<FC, BB4>::value;
  constexpr bool WEIGHT_FIRST = EQ && !SPLIT && std::is_same<FV, FC>::value;
  constexpr int NS = FOLD ? 4 : 2;  // source streams per table
  // The 64-bit weighted sums of BabyBear^4 weights on BabyBear values take
  // the registers of two more unrolled pairs.
  constexpr bool WIDE_SUMS = EQ && !WEIGHT_FIRST && std::is_same<FC, BB4>::value;
  // Narrow sources (at most 32 bytes a pair) are double-buffered, 32 bytes a
  // step: the next step's loads (tables and eq weights) are issued before this
  // step's arithmetic, so that every wave overlaps its own loads with its
  // compute; wider sources already fill the register budget and rely on the
  // other waves of the SIMD.
  constexpr int PAIR_BYTES = K * NS * (int)sizeof(S);
  constexpr bool PREFETCH = PAIR_BYTES <= 32;
  constexpr int U0 = PassUnroll<FS, K, FOLD>::value;
  // (SPLIT's nine lazy sums take the registers of a second prefetched pair.)
  constexpr int U = PREFETCH ? (SPLIT && EQ ? 1 : PF_STEP_BYTES / PAIR_BYTES)
                             : WIDE_SUMS && U0 > 2 ? 2 : U0;
  const size_t half = L >> 1;
  const size_t nxr_mask = ((size_t)1 << nxr_log) - 1;
  const bool factored = EQ && wl != nullptr && (stride & nxr_mask) == 0;
  WeightedSum<FC, FV> wsum[K + 1];
#pragma unroll
  for (int t = 0; t <= K; ++t) wsum[t].init();
  // SPLIT: csum[t][i] sums coefficient c_i of node t (weighted in eq mode).
  constexpr int NC = SPLIT ? K + 1 : 1;
  WeightedSum<FC, FS> csum[SPLIT ? K + 1 : 1][NC];
  typename FS::T bsum[SPLIT ? K + 1 : 1][NC];
#pragma unroll
  for (int t = 0; t < (SPLIT ? K + 1 : 1); ++t)
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      csum[t][i].init();
      bsum[t][i] = FS::zero();
    }
  // Eq weight of pair j: wsel[j >> wshift], times wr[j & wmask] when the
  // eq_r factor is not applied per thread (wmask 0 reads a dummy otherwise);
  // all uniform, so the weight loads take no branch.
  const T* wsel = wl ? wl : wr;
  const uint32_t wshift = wl ? nxr_log : 0u;
  const bool wmul = EQ && wl != nullptr && !factored;
  const size_t wmask = wmul ? nxr_mask : 0;
  // Sources of the U pairs starting at j0 into r; indices past the round are
  // clamped into it, so every load is issued whatever the step (an exact
  // count of loads in flight lets each wait leave the prefetched ones alone).
  auto load = [&](size_t j0, S (&r)[U][K][NS]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      size_t j = j0 + (size_t)u * stride;
      j = j < half ? j : half - 1;
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const S* s = src + (size_t)q * src_stride;
        r[u][q][0] = s[j];
        r[u][q][1] = s[j + half];
        if constexpr (FOLD) {
          r[u][q][2] = s[j + L];
          r[u][q][3] = s[j + half + L];
        }
      }
    }
  };
  // Fold (FOLD) and evaluate pair j from its sources r with weight w.
  // `fold`: whether the lazy sums fold after this pair (every second pair).
  auto pair = [&](size_t j, const S (&r)[K][NS], T w, bool fold) {
    if constexpr (SPLIT) {
      S a[K], da[K], b[K], db[K];  // node-0 values and steps of a and b
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const S s0 = r[q][0], s2 = r[q][1], s1 = r[q][2], s3 = r[q][3];
        const S e0 = FS::sub(s1, s0), e1 = FS::sub(s3, s2);
        T* d = dst + (size_t)q * dst_stride;
        d[j] = FC::add(Lift<FS, FC>::go(s0), FoldMul<FC, FS>::go(cm, e0));
        d[j + half] = FC::add(Lift<FS, FC>::go(s2), FoldMul<FC, FS>::go(cm, e1));
        a[q] = s0;
        da[q] = FS::sub(s2, s0);
        b[q] = e0;
        db[q] = FS::sub(e1, e0);
      }
#pragma unroll
      for (int t = 0; t <= K; ++t) {
        const S c0 = FS::mul(a[0], a[1]);
        const S c1 = FS::add(FS::mul(a[0], b[1]), FS::mul(b[0], a[1]));
        const S c2 = FS::mul(b[0], b[1]);
        if constexpr (EQ) {
          csum[t][0].add(w, c0, fold);
          csum[t][1].add(w, c1, fold);
          csum[t][2].add(w, c2, fold);
        } else {
          bsum[t][0] = FS::add(bsum[t][0], c0);
          bsum[t][1] = FS::add(bsum[t][1], c1);
          bsum[t][2] = FS::add(bsum[t][2], c2);
        }
        if (t < K) {
#pragma unroll
          for (int q = 0; q < K; ++q) {
            a[q] = FS::add(a[q], da[q]);
            b[q] = FS::add(b[q], db[q]);
          }
        }
      }
      __builtin_amdgcn_sched_barrier(0);
      return;
    }
    V lo[K], df[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      if constexpr (FOLD) {
        const S s0 = r[q][0], s2 = r[q][1], s1 = r[q][2], s3 = r[q][3];
        V a = FC::add(Lift<FS, FC>::go(s0), FoldMul<FC, FS>::go(cm, FS::sub(s1, s0)));
        V b = FC::add(Lift<FS, FC>::go(s2), FoldMul<FC, FS>::go(cm, FS::sub(s3, s2)));
        T* d = dst + (size_t)q * dst_stride;
        d[j] = a;
        d[j + half] = b;
        lo[q] = a;
        df[q] = FC::sub(b, a);
      } else {
        lo[q] = r[q][0];
        df[q] = FV::sub(r[q][1], r[q][0]);
      }
    }
    if constexpr (EQ && WEIGHT_FIRST) {
      V hi0 = FV::add(lo[0], df[0]);
      lo[0] = FC::mul(w, lo[0]);
      df[0] = FC::sub(FC::mul(w, hi0), lo[0]);
    }
#pragma unroll
    for (int t = 0; t <= K; ++t) {
      V prod = lo[0];
#pragma unroll
      for (int q = 1; q < K; ++q) prod = FV::mul(prod, lo[q]);
      if constexpr (EQ && !WEIGHT_FIRST) {
        wsum[t].add(w, prod, fold);
      } else {
        acc[t] = FC::add(acc[t], Lift<FV, FC>::go(prod));
      }
      if (t < K) {
#pragma unroll
        for (int q = 0; q < K; ++q) lo[q] = FV::add(lo[q], df[q]);
      }
    }
    // One pair's arithmetic at a time: interleaving the unrolled pairs would
    // only raise register pressure (their loads are already issued).
    __builtin_amdgcn_sched_barrier(0);
  };
  constexpr int UW = EQ ? U : 1;
  const size_t step = (size_t)U * stride;
  // Steps whose U pairs are in the round for every thread (start < stride).
  const size_t nfull = half / step;
  size_t j0 = start;
  S raw[U][K][NS];
  if (PREFETCH && nfull > 0) load(j0, raw);
  for (size_t k = 0; k < nfull; ++k, j0 += step) {
    // This step's weight operands, issued before the next step's sources so
    // that waiting for them leaves those in flight.
    T wa[UW], wb[UW];
    if constexpr (EQ) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const size_t j = j0 + (size_t)u * stride;
        wa[u] = wsel[j >> wshift];
        wb[u] = wr[j & wmask];
      }
    }
    S nxt[U][K][NS];
    if constexpr (PREFETCH) {
      load(j0 + step, nxt);
    } else {
      load(j0, raw);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      T w = FC::one();
      if constexpr (EQ) w = wmul ? FC::mul(wa[u], wb[u]) : wa[u];
      pair(j0 + (size_t)u * stride, raw[u], w, (u & 1) || u == U - 1);
    }
    if constexpr (PREFETCH) {
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int q = 0; q < K; ++q)
#pragma unroll
          for (int c = 0; c < NS; ++c) raw[u][q][c] = nxt[u][q][c];
    }
  }
  // The last, partial step: pairs j0 + u stride that are still in the round.
  if (j0 < half) {
    load(j0, raw);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t j = j0 + (size_t)u * stride;
      if (j < half) {
        T w = FC::one();
        if constexpr (EQ) {
          w = wsel[j >> wshift];
          if (wmul) w = FC::mul(w, wr[j & wmask]);
        }
        pair(j, raw[u], w, true);
      }
    }
  }
  if constexpr (SPLIT) {
    // node t = C0 + r C1 + r^2 C2.
    const T r1 = cm.r, r2 = FC::mul(cm.r, cm.r);
#pragma unroll
    for (int t = 0; t <= K; ++t) {
      T c[NC];
#pragma unroll
      for (int i = 0; i < NC; ++i)
        c[i] = EQ ? csum[t][i].get() : Lift<FS, FC>::go(bsum[t][i]);
      acc[t] = FC::add(acc[t], FC::add(c[0], FC::add(FC::mul(r1, c[1]), FC::mul(r2, c[2]))));
    }
  } else if constexpr (EQ && !WEIGHT_FIRST) {
#pragma unroll
    for (int t = 0; t <= K; ++t) acc[t] = FC::add(acc[t], wsum[t].get());
  }
  if (factored && start < half) {
    const T wrc = wr[start & nxr_mask];
#pragma unroll
    for (int t = 0; t <= K; ++t) acc[t] = FC::mul(acc[t], wrc);
  }
}

// --------------------------------------------------------- element bytes
// Raw challenge bytes -> canonical element (a squeezed word may be >= p; its
// value is the residue, reference zorch/transcript.py:97-119).
template <class F>
__device__ __forceinline__ typename F::T el_from_bytes(const uint8_t* b);
template <>
__device__ __forceinline__ uint32_t el_from_bytes<BB>(const uint8_t* b) {
  uint32_t w = (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) |
               ((uint32_t)b[3] << 24);
  return bb_canon(w);
}
template <>
__device__ __forceinline__ bb4 el_from_bytes<BB4>(const uint8_t* b) {
  bb4 r;
  for (int c = 0; c < 4; ++c) r.c[c] = el_from_bytes<BB>(b + 4 * c);
  return r;
}
template <>
__device__ __forceinline__ uint64_t el_from_bytes<GL>(const uint8_t* b) {
  uint64_t w = 0;
  for (int i = 0; i < 8; ++i) w |= (uint64_t)b[i] << (8 * i);
  return gl_canon(w);
}

// The same from 4-byte aligned little-endian words.
template <class F>
__device__ __forceinline__ typename F::T el_from_words(const uint32_t* w);
template <>
__device__ __forceinline__ uint32_t el_from_words<BB>(const uint32_t* w) {
  return bb_canon(w[0]);
}
template <>
__device__ __forceinline__ bb4 el_from_words<BB4>(const uint32_t* w) {
  return bb4_make(bb_canon(w[0]), bb_canon(w[1]), bb_canon(w[2]), bb_canon(w[3]));
}
template <>
__device__ __forceinline__ uint64_t el_from_words<GL>(const uint32_t* w) {
  return gl_canon(((uint64_t)w[1] << 32) | w[0]);
}

// A field element in vector registers (see vreg in hop.h).
template <class F>
__device__ __forceinline__ typename F::T vreg_el(typename F::T v);
template <>
__device__ __forceinline__ uint32_t vreg_el<BB>(uint32_t v) { return vreg(v); }
template <>
__device__ __forceinline__ bb4 vreg_el<BB4>(bb4 v) {
  return bb4_make(vreg(v.c[0]), vreg(v.c[1]), vreg(v.c[2]), vreg(v.c[3]));
}
template <>
__device__ __forceinline__ uint64_t vreg_el<GL>(uint64_t v) {
  return ((uint64_t)vreg((uint32_t)(v >> 32)) << 32) | vreg((uint32_t)v);
}

// LDS of the Fiat-Shamir hop: SHA-256 scratch and streaming state (midstate,
// pending bytes, length), the BLAKE3 streaming state, and the squeezed bytes.
struct HopLds {
  alignas(16) uint8_t pend[64];
  alignas(16) uint8_t blocks[HOP_MAX_BLOCKS * 64];
  alignas(16) uint32_t wk[HOP_MAX_BLOCKS * 64];
  alignas(16) uint8_t buf[HOP_MAX_MSG + 32];
  alignas(16) uint8_t out[64];
  uint32_t dgw[8];
  uint32_t h[8];
  uint32_t plen;
  uint64_t total;
  Blake3State b3;  // the whole stream state of a BLAKE3 transcript
};

#define ZS_LDS __attribute__((address_space(3)))

// The BLAKE3 hop out of line: its byte-serial stream code (every lane runs
// the same code on the LDS state) is large, and kept apart from the SHA-256
// path so that the latter stays small in the instruction cache.
static __device__ __attribute__((noinline)) void b3_fs_hop_ool(Blake3State* st, const uint8_t* m,
                                                               uint32_t msg_len, uint32_t item,
                                                               uint32_t limbs, uint8_t* out) {
  b3_fs_hop(*st, m, msg_len, item, limbs, out);
}

// observe(msg) + sample(limbs) on the transcript state in `s`; the squeezed
// bytes land in s.out. Wave 0, all lanes.
__device__ __forceinline__ void fs_hop_body(HopLds& s, const uint8_t* m, uint32_t msg_len,
                                            uint32_t item, uint32_t limbs, uint32_t kind) {
  if (kind == TR_BLAKE3) {
    b3_fs_hop_ool(&s.b3, m, msg_len, item, limbs, s.out);
    wave_sync();
    return;
  }
  WaveSha ws;
#pragma unroll
  for (int i = 0; i < 8; ++i) ws.h[i] = s.h[i];
  ws.plen = s.plen;
  ws.total = s.total;
  ws_fs_hop(ws, s.pend, s.buf, m, msg_len, item, limbs, s.blocks, s.wk, s.dgw, s.out);
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) s.h[i] = ws.h[i];
    s.plen = ws.plen;
    s.total = ws.total;
  }
  wave_sync();
}

// The hop out of line, for the pass and batch kernels: every such kernel of a
// translation unit calls this one copy of the two hashes' code (inlined into
// each of them it multiplied the code size and build time many times), and
// the LDS-typed arguments keep its accesses on LDS instructions. Callers all
// share one occupancy target, which the function inherits. A call waits for
// every outstanding memory operation, so the single-block tail kernel, which
// overlaps loads of the next round with the hop, inlines the body instead.
// Hop variants of finish_round: inlined, or the out-of-line copy of a group.
constexpr int HOP_INLINE = -1, HOP_PASS = 0, HOP_CHAIN = 1;

template <int GROUP>
static __device__ __attribute__((noinline)) void fs_hop(ZS_LDS HopLds* hl,
                                                        const ZS_LDS uint8_t* msg,
                                                        uint32_t msg_len, uint32_t item,
                                                        uint32_t limbs, uint32_t kind) {
  fs_hop_body(*(HopLds*)hl, (const uint8_t*)msg, msg_len, item, limbs, kind);
}

// LDS of the finisher: the hop's state and scratch, and the proof state
// between rounds (claim, challenge, eq mass), so that only the hop's own
// working set occupies registers.
template <class FC>
struct FinishLds : HopLds {
  alignas(16) typename FC::T msg[MAXPTS];
  typename FC::T claim, r, eq_prev;
  // Round constants, read once per kernel: nodes, extension coefficients,
  // Lagrange denominators; the current round's eq line coefficients.
  typename FC::T nodes[MAXPTS], ext_coef[MAXPTS], lag_inv[MAXPTS];
  typename FC::T eqA, eqB;
  // The wire form's maps (Ctl::wire ...); snat holds the natural values.
  uint32_t wire, mlen, ncoef, claim_every;
  typename FC::T wmat[MAXPTS][MAXPTS], cmat[MAXPTS][MAXPTS], cclaim[MAXPTS];
  alignas(16) typename FC::T snat[MAXPTS];
  // Every round's eq line coefficients, for finishers that run many rounds
  // (fin_eq_table): a round reads them from LDS instead of memory.
  typename FC::T eqab[2 * 32];
};

// Control block -> finisher LDS (wave 0, all lanes).
template <class FC>
__device__ __forceinline__ void fin_load(FinishLds<FC>& sm, const Ctl<FC>* ctl) {
  const uint32_t lane = threadIdx.x & 63;
  if (ctl->tr_kind == TR_BLAKE3) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&ctl->tr.b3);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&sm.b3);
    for (uint32_t i = lane; i < sizeof(Blake3State) / 4; i += 64u) dst[i] = src[i];
  } else {
    sm.pend[lane] = ctl->tr.sha.pending[lane];
    if (lane < 8u) sm.h[lane] = ctl->tr.sha.h[lane];
  }
  if (lane == 0) {
    if (ctl->tr_kind != TR_BLAKE3) {
      sm.plen = (uint32_t)ctl->tr.sha.pending_len;
      sm.total = ctl->tr.sha.total_len;
    }
    sm.claim = ctl->claim;
    sm.r = ctl->r;
    sm.eq_prev = ctl->eq_prev;
  }
  if (lane < (uint32_t)MAXPTS) {
    sm.nodes[lane] = ctl->nodes[lane];
    sm.ext_coef[lane] = ctl->ext_coef[lane];
    sm.lag_inv[lane] = ctl->lag_inv[lane];
    sm.cclaim[lane] = ctl->cclaim[lane];
  }
  if (lane == 0) {
    sm.wire = ctl->wire;
    sm.mlen = ctl->mlen;
    sm.ncoef = ctl->ncoef;
    sm.claim_every = ctl->claim_every;
  }
  if (ctl->wire != WIRE_VALUES) {
    for (uint32_t i = lane; i < (uint32_t)(MAXPTS * MAXPTS); i += 64u) {
      (&sm.wmat[0][0])[i] = (&ctl->wmat[0][0])[i];
      (&sm.cmat[0][0])[i] = (&ctl->cmat[0][0])[i];
    }
  }
  wave_sync();
}

// The eq line coefficients of `round` -> finisher LDS (wave 0, lane 0); read
// by finish_round after the next wave or block barrier.
template <class FS, class FC, bool EQ>
__device__ __forceinline__ void fin_eq(FinishLds<FC>& sm, const StageArgs<FS, FC>& a,
                                       uint32_t round) {
  if (EQ && (threadIdx.x & 63) == 0) {
    sm.eqA = a.eq_ab[2 * round];
    sm.eqB = a.eq_ab[2 * round + 1];
  }
}

// Every round's eq line coefficients -> finisher LDS (wave 0, all lanes), for
// fin_eq_lds.
template <class FS, class FC, bool EQ>
__device__ __forceinline__ void fin_eq_table(FinishLds<FC>& sm, const StageArgs<FS, FC>& a) {
  if constexpr (EQ) {
    const uint32_t lane = threadIdx.x & 63;
    if (lane < 2 * a.n) sm.eqab[lane] = a.eq_ab[lane];
    wave_sync();
  }
}

// The eq line coefficients of `round` from the table fin_eq_table loaded.
template <class FC, bool EQ>
__device__ __forceinline__ void fin_eq_lds(FinishLds<FC>& sm, uint32_t round) {
  if (EQ && (threadIdx.x & 63) == 0) {
    sm.eqA = sm.eqab[2 * round];
    sm.eqB = sm.eqab[2 * round + 1];
  }
}

// Finisher LDS -> control block (wave 0, all lanes); pending bytes past the
// pending length are stored as zero.
template <class FC>
__device__ __forceinline__ void fin_store(const FinishLds<FC>& sm, Ctl<FC>* ctl) {
  const uint32_t lane = threadIdx.x & 63;
  if (ctl->tr_kind == TR_BLAKE3) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&sm.b3);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&ctl->tr.b3);
    for (uint32_t i = lane; i < sizeof(Blake3State) / 4; i += 64u) dst[i] = src[i];
    if (lane == 0) {
      ctl->claim = sm.claim;
      ctl->r = sm.r;
      ctl->eq_prev = sm.eq_prev;
    }
    return;
  }
  const uint32_t plen = sm.plen;
  ctl->tr.sha.pending[lane] = lane < plen ? sm.pend[lane] : 0;
  if (lane < 8u) ctl->tr.sha.h[lane] = sm.h[lane];
  if (lane == 0) {
    ctl->tr.sha.pending_len = plen;
    ctl->tr.sha.total_len = sm.total;
    ctl->claim = sm.claim;
    ctl->r = sm.r;
    ctl->eq_prev = sm.eq_prev;
  }
}

// The claim a round reduces to at r from its message `msg` (LDS, the wire
// form's elements) and the claim before it; every lane of the wave gets it.
// Values on the wire: s(r) = sum_t s(t) L_t(r), L_t(r) = lag_inv[t]
// prod_{j != t} (r - j), lane t (of each eight-lane group) forming term t.
// Another wire form: its reduction map, lane k forming c_k r^k.
template <class FC, int NPTS>
__device__ __forceinline__ typename FC::T round_claim(const FinishLds<FC>& sm,
                                                      const typename FC::T* msg,
                                                      typename FC::T r, typename FC::T claim) {
  using T = typename FC::T;
  const uint32_t t = threadIdx.x & 7u;
  const uint32_t tn = t < (uint32_t)NPTS ? t : (uint32_t)(NPTS - 1);
  if (sm.wire == WIRE_VALUES) {
    T basis = sm.lag_inv[tn];
#pragma unroll
    for (int j = 0; j < NPTS; ++j) {
      const T d = FC::sub(r, sm.nodes[j]);
      basis = FC::mul(basis, t == (uint32_t)j ? FC::one() : d);
    }
    const T m = t < (uint32_t)NPTS ? msg[tn] : FC::zero();
    return lanes8_sum<FC>(FC::mul(m, basis));
  }
  const uint32_t kk = t < sm.ncoef ? t : 0u;
  T c = FC::mul(sm.cclaim[kk], claim);
  for (uint32_t j = 0; j < sm.mlen; ++j) c = FC::add(c, FC::mul(sm.cmat[kk][j], msg[j]));
  T rk = FC::one();
  for (uint32_t i = 0; i < kk; ++i) rk = FC::mul(rk, r);
  return lanes8_sum<FC>(t < sm.ncoef ? FC::mul(c, rk) : FC::zero());
}

// Wave 0 only. Lane t = lane & 7 works on node t (the eight-lane groups of
// the wave agree): message from the node sums, Fiat-Shamir hop, bind, eq-mass
// update, the claim reduction when `need_claim`, and (last round) the final
// fold. The proof state is read from and written back to `sm`.
template <class FS, class FC, int K, bool EQ, int HOP>
__device__ __forceinline__ void finish_round(const StageArgs<FS, FC>& a, uint32_t round,
                                             bool need_claim, const typename FC::T* sums,
                                             FinishLds<FC>& sm,
                                             const typename FC::T* last_lo_hi_src,
                                             const typename FS::T* last_in, size_t last_stride,
                                             bool last_src_is_input) {
  using T = typename FC::T;
  constexpr int NPTS = K + 1 + (EQ ? 1 : 0);
  constexpr int EB = sizeof(T);
  static_assert(NPTS <= 8, "one node per lane of an eight-lane group");
  static_assert(NPTS * EB <= (int)HOP_MAX_MSG, "hop message bound");
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t t = lane & 7;
  const uint32_t tn = t < (uint32_t)NPTS ? t : (uint32_t)(NPTS - 1);  // in-bounds node index
  const Ctl<FC>* ctl = a.ctl;
  {
    T tv = sums[t <= (uint32_t)K ? t : (uint32_t)K];
    if (t > (uint32_t)K) tv = FC::zero();
    T msg;
    if constexpr (EQ) {
      const T A = sm.eqA, B = sm.eqB;
      const T node = sm.nodes[tn];
      const T coef = sm.ext_coef[t <= (uint32_t)K ? t : 0u];
      // Degree-K extension to node K+1: sum_i (-1)^(K-i) C(K+1, i) t(i).
      const T ext = lanes8_sum<FC>(t <= (uint32_t)K ? FC::mul(tv, coef) : FC::zero());
      if (t == (uint32_t)(K + 1)) tv = ext;
      // l(t) = eq_prev ((1 - w) + t (2 w - 1)); lanes past the nodes carry zero.
      msg = FC::mul(FC::mul(sm.eq_prev, FC::add(B, FC::mul(node, A))), tv);
    } else {
      msg = tv;
    }
    if (sm.wire == WIRE_VALUES) {
      if (lane < (uint32_t)NPTS) {
        a.msgs[(size_t)round * NPTS + lane] = msg;
        sm.msg[lane] = msg;
      }
    } else {
      // Another wire form: element j of the message is row j of the wire
      // map applied to the natural values.
      if (lane < (uint32_t)NPTS) sm.snat[lane] = msg;
      wave_sync();
      const uint32_t j = t < sm.mlen ? t : 0u;
      T m = FC::zero();
#pragma unroll
      for (int i = 0; i < NPTS; ++i) m = FC::add(m, FC::mul(sm.wmat[j][i], sm.snat[i]));
      if (lane < sm.mlen) {
        a.msgs[(size_t)round * NPTS + lane] = m;
        sm.msg[lane] = m;
      }
    }
  }
  wave_sync();
  if (lane == 0) trace_mark(a.trace, round, 4);
  static_assert(NPTS * EB + 32 <= (int)sizeof(HopLds::buf), "hop framing buffer");
  const uint32_t mlen = sm.wire == WIRE_VALUES ? (uint32_t)NPTS : sm.mlen;
  if constexpr (HOP == HOP_INLINE) {
    fs_hop_body(sm, reinterpret_cast<const uint8_t*>(sm.msg), mlen * (uint32_t)EB,
                ctl->item_bytes, ctl->limbs, ctl->tr_kind);
  } else {
    fs_hop<HOP>((ZS_LDS HopLds*)static_cast<HopLds*>(&sm), (const ZS_LDS uint8_t*)sm.msg,
                mlen * (uint32_t)EB, ctl->item_bytes, ctl->limbs, ctl->tr_kind);
  }
  if (lane == 0) trace_mark(a.trace, round, 5);
  if (lane < (uint32_t)EB) a.point[(size_t)round * EB + lane] = sm.out[lane];
  const T r = el_from_words<FC>(reinterpret_cast<const uint32_t*>(sm.out));
  T eq_next = FC::zero();
  if constexpr (EQ) eq_next = FC::mul(sm.eq_prev, FC::add(sm.eqB, FC::mul(r, sm.eqA)));
  T claim = FC::zero();
  need_claim = need_claim || sm.claim_every != 0;
  if (need_claim) claim = round_claim<FC, NPTS>(sm, sm.msg, r, sm.claim);
  wave_sync();
  if (lane == 0) {
    sm.r = r;
    if (need_claim) sm.claim = claim;
    if constexpr (EQ) sm.eq_prev = eq_next;
    trace_mark(a.trace, round, 6);
  }
  if (round + 1 == a.n && a.final_out && lane < a.k) {
    const uint32_t q = lane;
    T lo, hi;
    if (last_src_is_input) {
      lo = Lift<FS, FC>::go(last_in[(size_t)q * last_stride]);
      hi = Lift<FS, FC>::go(last_in[(size_t)q * last_stride + 1]);
    } else {
      lo = last_lo_hi_src[(size_t)q * last_stride];
      hi = last_lo_hi_src[(size_t)q * last_stride + 1];
    }
    a.final_out[q] = FC::add(lo, FC::mul(r, FC::sub(hi, lo)));
    if (EQ && q == 0) a.final_out[a.k] = eq_next;
  }
  wave_sync();
}

// ------------------------------------------------------------- kernels
template <class FC, int NA>
struct RoundLds {
  typename FC::T lds[(PASS_THREADS / 64) * NA];
  typename FC::T out[NA];
  uint64_t sum64[PASS_THREADS / SLOTW][SLOTW];
  FinishLds<FC> sm;
  uint32_t is_last;
};

// LDS of a pass kernel: the round's LDS and a scratch copy of the hop state
// for the finisher block's warm-up run.
template <class FC, int NA>
struct PassLds {
  RoundLds<FC, NA> round;
  HopLds warm;
};

// Add a block's node sums (NA elements in LDS), word by word, to slot row
// blk % NSLOT.
template <class FC, int NA>
__device__ __forceinline__ void slot_row_add(uint64_t* slots, uint32_t blk,
                                             const typename FC::T* out) {
  constexpr int W = SlotOps<FC>::W;
  constexpr int NW = NA * W;
  static_assert(NW <= SLOTW && NSLOT % (PASS_THREADS / SLOTW) == 0, "slot row layout");
  if (threadIdx.x < NW) {
    uint64_t* slot = slots + (size_t)(blk & (NSLOT - 1)) * SLOTW + threadIdx.x;
    const uint32_t word = SlotOps<FC>::word(out[threadIdx.x / W], threadIdx.x % W);
    __hip_atomic_fetch_add(slot, (uint64_t)word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Sum of the slot rows into sh.out (NA elements; sh a RoundLds), which it
// clears for the next round; ends with a barrier. Every block that added to
// the rows has completed its additions.
template <class FC, int NA, class Lds>
__device__ __forceinline__ void slot_rows_sum(uint64_t* slots, Lds& sh) {
  constexpr int W = SlotOps<FC>::W;
  constexpr int NW = NA * W;
  constexpr int G = PASS_THREADS / SLOTW;
  // Thread (g, w) sums word w of rows g, g + G, ...
  const uint32_t w = threadIdx.x % SLOTW, g = threadIdx.x / SLOTW;
  uint64_t s = 0;
  if (w < (uint32_t)NW) {
#pragma unroll
    for (int i = 0; i < NSLOT / G; ++i)
      s += __hip_atomic_load(slots + (size_t)(g + i * G) * SLOTW + w, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
  }
  sh.sum64[g][w] = s;
  __syncthreads();
  // Every row has been read: clear them for the next round.
  if (w < (uint32_t)NW) {
#pragma unroll
    for (int i = 0; i < NSLOT / G; ++i)
      __hip_atomic_store(slots + (size_t)(g + i * G) * SLOTW + w, (uint64_t)0, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  }
  if (threadIdx.x < NA) {
    uint64_t tot[W];
#pragma unroll
    for (int c = 0; c < W; ++c) {
      uint64_t v = 0;
#pragma unroll
      for (int q = 0; q < G; ++q) v += sh.sum64[q][threadIdx.x * W + c];
      tot[c] = v;
    }
    sh.out[threadIdx.x] = SlotOps<FC>::join(tot);
  }
  __syncthreads();
}

// One block's share of a multi-block round: block `blk` of `nblk` accumulates
// its pairs and adds its node sums into slot row blk % NSLOT; the last block
// of the round to take a ticket sums the rows, clears them and finishes the
// round.
template <class FT, class FS, class FC, int K, bool EQ, bool FOLD>
__device__ __forceinline__ void round_block(const StageArgs<FT, FC>& a, const typename FS::T* src,
                                            size_t src_stride, uint32_t L, uint32_t round,
                                            bool need_claim, uint32_t blk, uint32_t nblk,
                                            RoundLds<FC, K + 1>& sh) {
  using T = typename FC::T;
  constexpr int NA = K + 1;
  T acc[NA];
#pragma unroll
  for (int t = 0; t < NA; ++t) acc[t] = FC::zero();
  const T* wl = nullptr;
  const T* wr = nullptr;
  uint32_t nxr_log = 0;
  if constexpr (EQ) select_eq<FC>(round, a.n, a.eq_l_fam, a.eq_r_fam, &a.ctl->one, wl, wr, nxr_log);
  ChalMul<FC> cm;
  if constexpr (FOLD) cm.init(a.ctl->r);
  if (blk == 0 && threadIdx.x == 0) trace_mark(a.trace, round, 1);
  const size_t start = (size_t)blk * PASS_THREADS + threadIdx.x;
  const size_t stride = (size_t)nblk * PASS_THREADS;
  pass_accumulate<FS, FC, K, EQ, FOLD>(src, src_stride, a.work, a.work_stride, L, cm, wl, wr,
                                       nxr_log, start, stride, acc);
  block_reduce<FC, NA, PASS_THREADS>(acc, sh.lds, sh.out);
  if (nblk > 1) {
    slot_row_add<FC, NA>(a.slots, blk, sh.out);
    // The slot additions complete before the ticket is taken.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t prev =
          __hip_atomic_fetch_add(&a.ctl->counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      sh.is_last = (prev == nblk - 1) ? 1u : 0u;
    }
    __syncthreads();
    if (!sh.is_last) return;
    if (threadIdx.x == 0) trace_mark(a.trace, round, 2);
    if (threadIdx.x < 64) {
      fin_eq<FT, FC, EQ>(sh.sm, a, round);
      fin_load(sh.sm, a.ctl);
    }
    slot_rows_sum<FC, NA>(a.slots, sh);
  } else if (threadIdx.x < 64) {
    fin_eq<FT, FC, EQ>(sh.sm, a, round);
    fin_load(sh.sm, a.ctl);
  }
  if (threadIdx.x < 64) {
    finish_round<FT, FC, K, EQ, HOP_PASS>(a, round, need_claim, sh.out, sh.sm, a.work, a.in,
                                          round == 0 ? a.in_stride : a.work_stride, round == 0);
    fin_store(sh.sm, a.ctl);
  }
  if (nblk > 1 && threadIdx.x == 0)
    __hip_atomic_store(&a.ctl->counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Multi-block pass of one round of one proof. The streaming loop needs fewer
// registers than the finisher that one block runs at the end; PASS_WAVES_PER_EU
// keeps the occupancy of the loop (the finisher may spill a few registers):
// four waves a SIMD, what a grid of 1024 blocks of 256 threads places on 256
// compute units.
#define PASS_WAVES_PER_EU __attribute__((amdgpu_waves_per_eu(4, 8)))

// Polls of a device wait, each after a short sleep: far above any real wait,
// the limit only keeps a fault from hanging the device. A wait that reaches
// it marks the proof faulted (the host reports an error) and the block exits.
constexpr uint32_t CHAIN_SPIN_LIMIT = 1u << 22;

// Wait until *p >= target, then acquire (device scope). False on a fault.
// The polls are relaxed: an acquiring load invalidates the caches it passes
// through, and a storm of them stalls every other wave on the device.
__device__ __forceinline__ bool chain_wait(uint32_t* p, uint32_t target, uint32_t* fault) {
  for (uint32_t i = 0; i < CHAIN_SPIN_LIMIT; ++i) {
    if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      return true;
    }
    if (__hip_atomic_load(fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
    __builtin_amdgcn_s_sleep(2);
  }
  __hip_atomic_store(fault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return false;
}

// One round as 1 + nblk blocks: blocks 1..nblk stream the round's pairs, add
// their node sums to the slot rows and arrive at the control block's counter
// (release); block 0 takes no pairs. While the others stream, block 0 loads
// the proof state and round constants and runs the hop once on a scratch
// copy of the transcript state (the same code path as the real hop, whose
// length depends only on the pending length and the message size), so that
// the finisher's code sits in its compute unit's instruction cache; then it
// waits for every arrival, sums and clears the rows and finishes the round.
// Only block 0 waits, on blocks that never wait, so residency is not needed.
template <class FT, class FS, class FC, int K, bool EQ, bool FOLD>
__global__ __launch_bounds__(PASS_THREADS) PASS_WAVES_PER_EU void pass_kernel(StageArgs<FT, FC> a,
                                                            const typename FS::T* src,
                                                            size_t src_stride, uint32_t L,
                                                            uint32_t round, uint32_t need_claim) {
  using T = typename FC::T;
  constexpr int NA = K + 1;
  constexpr int NPTS = K + 1 + (EQ ? 1 : 0);
  __shared__ PassLds<FC, NA> ps;
  RoundLds<FC, NA>& sh = ps.round;
  const uint32_t tid = threadIdx.x, nblk = gridDim.x - 1;
  Ctl<FC>* ctl = a.ctl;
  if (blockIdx.x == 0) {
    if (tid < 64) {
      fin_eq<FT, FC, EQ>(sh.sm, a, round);
      fin_load(sh.sm, ctl);
      // Warm-up: the hop on a copy of the stream state; the message is
      // whatever the (not yet written) message buffer holds.
      const uint32_t* from = reinterpret_cast<const uint32_t*>(static_cast<HopLds*>(&sh.sm));
      uint32_t* to = reinterpret_cast<uint32_t*>(&ps.warm);
      for (uint32_t i = tid; i < sizeof(HopLds) / 4; i += 64u) to[i] = from[i];
      wave_sync();
      fs_hop<HOP_PASS>((ZS_LDS HopLds*)&ps.warm, (const ZS_LDS uint8_t*)sh.sm.msg,
                       (uint32_t)(NPTS * sizeof(T)), ctl->item_bytes, ctl->limbs, ctl->tr_kind);
    }
    if (tid == 0) sh.is_last = chain_wait(&ctl->counter, nblk, &ctl->fault) ? 1u : 0u;
    __syncthreads();
    if (!sh.is_last) return;
    if (tid == 0) trace_mark(a.trace, round, 2);
    slot_rows_sum<FC, NA>(a.slots, sh);
    if (tid < 64) {
      finish_round<FT, FC, K, EQ, HOP_PASS>(a, round, need_claim != 0, sh.out, sh.sm, a.work,
                                            a.in, round == 0 ? a.in_stride : a.work_stride,
                                            round == 0);
      fin_store(sh.sm, ctl);
    }
    if (tid == 0) __hip_atomic_store(&ctl->counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  const uint32_t blk = blockIdx.x - 1;
  T acc[NA];
#pragma unroll
  for (int t = 0; t < NA; ++t) acc[t] = FC::zero();
  const T* wl = nullptr;
  const T* wr = nullptr;
  uint32_t nxr_log = 0;
  if constexpr (EQ) select_eq<FC>(round, a.n, a.eq_l_fam, a.eq_r_fam, &ctl->one, wl, wr, nxr_log);
  ChalMul<FC> cm;
  if constexpr (FOLD) cm.init(ctl->r);
  if (blk == 0 && tid == 0) trace_mark(a.trace, round, 1);
  pass_accumulate<FS, FC, K, EQ, FOLD>(src, src_stride, a.work, a.work_stride, L, cm, wl, wr,
                                       nxr_log, (size_t)blk * PASS_THREADS + tid,
                                       (size_t)nblk * PASS_THREADS, acc);
  block_reduce<FC, NA, PASS_THREADS>(acc, sh.lds, sh.out);
  slot_row_add<FC, NA>(a.slots, blk, sh.out);
  // The slot additions (device-scope atomics) are complete before the
  // arrival. The folded entries reach the next round through the kernel
  // boundary; only the final fold of the last round reads them in this
  // kernel, which takes a releasing arrival (a releasing atomic writes back
  // the whole L2 of its die, too costly for every block of every round).
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (tid == 0) {
    if (round + 1 == a.n)
      __hip_atomic_fetch_add(&ctl->counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    else
      __hip_atomic_fetch_add(&ctl->counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ------------------------------------------------------------ chained rounds

// Blocks that share round `round` of a chained launch of `grid` blocks: one
// per PASS_THREADS pairs, at most the grid. Non-increasing in the round, so
// a block without work in a round has none later.
__device__ __forceinline__ uint32_t chain_blocks(uint32_t n, uint32_t round, uint32_t grid) {
  const uint32_t per = (1u << (n - round - 1)) / PASS_THREADS;
  return per < 1u ? 1u : per < grid ? per : grid;
}

// Rounds [first, last) (all >= 2: tables in the challenge field) of one proof
// in one persistent launch, whose grid is small enough for all its blocks to
// be resident at once (blocks wait for one another). Per round, every block with work folds and evaluates its
// grid-stride share, adds its node sums to the slot rows and arrives at the
// control block's counter (release). Block 0 waits for every arrival, sums
// and clears the rows, runs the finisher (the proof state stays in its LDS
// across rounds), publishes the challenge and bumps `done` (release); the
// other blocks wait for `done` (acquire) and read the challenge. Against one
// launch per round this removes the launch gap and keeps the finisher's code
// warm in one compute unit's instruction cache. The claim is reduced at
// round claim_last - 1 only.
template <class FT, class FC, int K, bool EQ>
__global__ __launch_bounds__(PASS_THREADS) void chain_kernel(
    StageArgs<FT, FC> a, uint32_t first, uint32_t last, uint32_t claim_last) {
  using T = typename FC::T;
  constexpr int NA = K + 1;
  __shared__ RoundLds<FC, NA> sh;
  __shared__ T r_sh;
  __shared__ uint32_t ok_sh;
  const uint32_t blk = blockIdx.x, tid = threadIdx.x;
  Ctl<FC>* ctl = a.ctl;
  if (blk == 0 && tid == 0) trace_mark(a.trace, first, 7);
  if (blk == 0 && tid < 64) fin_load(sh.sm, ctl);
  if (tid == 0) r_sh = ctl->r;
  __syncthreads();
  if (blk == 0 && tid == 0) trace_mark(a.trace, first, 8);
  uint32_t arrivals = 0;  // counter value once every block has arrived in this round
  for (uint32_t round = first; round < last; ++round) {
    const uint32_t nb = chain_blocks(a.n, round, gridDim.x);
    if (blk >= nb) break;
    if (nb > 1) arrivals += nb;
    T acc[NA];
#pragma unroll
    for (int t = 0; t < NA; ++t) acc[t] = FC::zero();
    const T* wl = nullptr;
    const T* wr = nullptr;
    uint32_t nxr_log = 0;
    if constexpr (EQ) select_eq<FC>(round, a.n, a.eq_l_fam, a.eq_r_fam, &ctl->one, wl, wr, nxr_log);
    ChalMul<FC> cm;
    cm.init(r_sh);
    if (blk == 0 && tid < 64) fin_eq<FT, FC, EQ>(sh.sm, a, round);
    if (blk == 0 && tid == 0) trace_mark(a.trace, round, 1);
    pass_accumulate<FC, FC, K, EQ, true>(a.work, a.work_stride, a.work, a.work_stride,
                                         1u << (a.n - round), cm, wl, wr, nxr_log,
                                         (size_t)blk * PASS_THREADS + tid,
                                         (size_t)nb * PASS_THREADS, acc);
    block_reduce<FC, NA, PASS_THREADS>(acc, sh.lds, sh.out);
    if (nb > 1) {
      slot_row_add<FC, NA>(a.slots, blk, sh.out);
      // Slot additions and folded entries are complete before the arrival.
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (tid == 0) __hip_atomic_fetch_add(&ctl->counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (blk == 0) {
      if (nb > 1) {
        if (tid == 0) ok_sh = chain_wait(&ctl->counter, arrivals, &ctl->fault) ? 1u : 0u;
        __syncthreads();
        if (!ok_sh) return;
        if (tid == 0) trace_mark(a.trace, round, 2);
        slot_rows_sum<FC, NA>(a.slots, sh);
      }
      if (tid < 64) {
        finish_round<FT, FC, K, EQ, HOP_CHAIN>(a, round, round + 1 == claim_last, sh.out, sh.sm,
                                               a.work, a.in, a.work_stride, false);
      }
      if (tid == 0) {
        r_sh = sh.sm.r;
        ctl->r = sh.sm.r;
        __hip_atomic_store(&ctl->done, round - first + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
      __syncthreads();
    } else if (round + 1 < last && blk < chain_blocks(a.n, round + 1, gridDim.x)) {
      if (tid == 0) {
        const bool ok = chain_wait(&ctl->done, round - first + 1, &ctl->fault);
        if (ok) r_sh = ctl->r;
        ok_sh = ok ? 1u : 0u;
      }
      __syncthreads();
      if (!ok_sh) return;
    }
  }
  if (blk == 0) {
    // Every other block has arrived for the last time and waits no more.
    if (tid < 64) fin_store(sh.sm, ctl);
    if (tid == 0) {
      __hip_atomic_store(&ctl->counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ctl->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// LDS of the single-block tail loop.
template <class FC, int NA, int THREADS>
struct TailLds {
  typename FC::T lds[(THREADS / 64) * NA];
  typename FC::T out[NA];
  typename FC::T lastv[MAXK][2];  // the last round's folded pair, for the final fold
  FinishLds<FC> sm;
};

// ------------------------------------------------- one-pair-per-thread rounds
// A round r >= 2 with at most one pair per thread: thread j folds and
// evaluates pair j alone. Its sources are table q's entries
// T[j], T[j + L/2], T[j + L], T[j + L/2 + L] of the previous round's tables
// (length 2L, in place in `work`), and its eq weight is wa * wb.
template <class FC, int K>
struct PairSrc {
  typename FC::T s[K][4];
  typename FC::T wa, wb;
};

template <class F>
__device__ __forceinline__ typename F::T shfl_el(typename F::T v, int src);
template <>
__device__ __forceinline__ uint32_t shfl_el<BB>(uint32_t v, int src) {
  return __shfl(v, src, 64);
}
template <>
__device__ __forceinline__ bb4 shfl_el<BB4>(bb4 v, int src) {
  return bb4_make(__shfl(v.c[0], src, 64), __shfl(v.c[1], src, 64), __shfl(v.c[2], src, 64),
                  __shfl(v.c[3], src, 64));
}
template <>
__device__ __forceinline__ uint64_t shfl_el<GL>(uint64_t v, int src) {
  uint32_t lo = __shfl((uint32_t)v, src, 64), hi = __shfl((uint32_t)(v >> 32), src, 64);
  return ((uint64_t)hi << 32) | lo;
}

template <class FS, class FC, int K, bool EQ>
__device__ __forceinline__ void pair_weight_load(const StageArgs<FS, FC>& a, uint32_t round,
                                                 uint32_t j, PairSrc<FC, K>& ps) {
  if constexpr (EQ) {
    const typename FC::T* wl = nullptr;
    const typename FC::T* wr = nullptr;
    uint32_t nxr_log = 0;
    select_eq<FC>(round, a.n, a.eq_l_fam, a.eq_r_fam, &a.ctl->one, wl, wr, nxr_log);
    if (wl) {
      ps.wa = wl[j >> nxr_log];
      ps.wb = wr[j & ((1u << nxr_log) - 1u)];
    } else {
      ps.wa = wr[j];
      ps.wb = FC::one();
    }
  }
}

template <class FS, class FC, int K, bool EQ>
__device__ __forceinline__ void pair_load(const StageArgs<FS, FC>& a, uint32_t round, uint32_t j,
                                          PairSrc<FC, K>& ps) {
  const size_t L = (size_t)1 << (a.n - round), half = L >> 1;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const typename FC::T* s = a.work + (size_t)q * a.work_stride;
    ps.s[q][0] = s[j];
    ps.s[q][1] = s[j + half];
    ps.s[q][2] = s[j + L];
    ps.s[q][3] = s[j + half + L];
  }
  pair_weight_load<FS, FC, K, EQ>(a, round, j, ps);
}

// Fold pair j's sources at the previous challenge into av = T'[j] and
// bv = T'[j + L/2] (stored to `work` as well), and add the pair's node values
// to acc: weight-first in eq mode, as in pass_accumulate.
template <class FS, class FC, int K, bool EQ>
__device__ __forceinline__ void pair_fold_eval(const StageArgs<FS, FC>& a, uint32_t round,
                                               uint32_t j, const PairSrc<FC, K>& ps,
                                               const ChalMul<FC>& cm, typename FC::T (&av)[K],
                                               typename FC::T (&bv)[K],
                                               typename FC::T (&acc)[K + 1]) {
  using T = typename FC::T;
  const size_t half = (size_t)1 << (a.n - round - 1);
  T lo[K], df[K];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const T x = FC::add(ps.s[q][0], cm.mul(FC::sub(ps.s[q][2], ps.s[q][0])));
    const T y = FC::add(ps.s[q][1], cm.mul(FC::sub(ps.s[q][3], ps.s[q][1])));
    T* d = a.work + (size_t)q * a.work_stride;
    d[j] = x;
    d[j + half] = y;
    av[q] = x;
    bv[q] = y;
    lo[q] = x;
    df[q] = FC::sub(y, x);
  }
  if constexpr (EQ) {
    const T w = FC::mul(ps.wa, ps.wb);
    const T hi0 = FC::add(lo[0], df[0]);
    lo[0] = FC::mul(w, lo[0]);
    df[0] = FC::sub(FC::mul(w, hi0), lo[0]);
  }
#pragma unroll
  for (int t = 0; t <= K; ++t) {
    T prod = lo[0];
#pragma unroll
    for (int q = 1; q < K; ++q) prod = FC::mul(prod, lo[q]);
    acc[t] = FC::add(acc[t], prod);
    if (t < K) {
#pragma unroll
      for (int q = 0; q < K; ++q) lo[q] = FC::add(lo[q], df[q]);
    }
  }
}

// The small tail rounds [first, last) of one proof in one block of THREADS
// threads, in three modes:
//  * rounds that read the input tables or have more pairs than threads are
//    block-stride passes;
//  * later rounds with more than 64 pairs give each thread at most one pair;
//    in the tail kernel (whose hop is inlined) the next round's sources,
//    written by this round and visible after the reduction's barriers, are
//    loaded into registers while wave 0 runs the hop;
//  * once a round has at most 64 pairs, wave 0 runs the remaining rounds
//    alone: each lane keeps its folded pair in registers, the next round's
//    sources come from a lane shuffle, the node sums from a wave reduction,
//    and no block barrier is taken (the other waves wait at the end).
// One finisher call site serves all three. The transcript, challenge and eq
// mass stay in LDS until the end.
template <class FS, class FC, int K, bool EQ, int THREADS>
__device__ __forceinline__ void tail_rounds(const StageArgs<FS, FC>& a, uint32_t first,
                                            uint32_t last, TailLds<FC, K + 1, THREADS>& sh) {
  using T = typename FC::T;
  constexpr int NA = K + 1;
  constexpr int HOP = THREADS == TAIL_THREADS ? HOP_INLINE : HOP_PASS;
  const uint32_t tid = threadIdx.x, lane = tid & 63u;
  if (tid < 64) fin_load(sh.sm, a.ctl);
  __syncthreads();
  PairSrc<FC, K> ps;
  T av[K], bv[K];
#pragma unroll
  for (int q = 0; q < K; ++q) av[q] = bv[q] = FC::zero();
  bool have_src = false;  // ps holds this round's sources (prefetched)
  bool from_regs = false; // this round's sources are the previous round's folded pairs
  for (uint32_t round = first; round < last; ++round) {
    const uint32_t P = 1u << (a.n - round - 1);
    const bool stride_pass = round < 2 || P > (uint32_t)THREADS;
    const bool wave_mode = !stride_pass && P <= 64u;
    // Wave mode lasts to the end: the other waves go straight to the final barrier.
    if (wave_mode && tid >= 64) break;
    ChalMul<FC> cm;
    if (round > 0) cm.init(sh.sm.r);
    if (tid < 64) fin_eq<FS, FC, EQ>(sh.sm, a, round);
    T acc[NA];
#pragma unroll
    for (int t = 0; t < NA; ++t) acc[t] = FC::zero();
    if (tid == 0) trace_mark(a.trace, round, 1);
    if (stride_pass) {
      const uint32_t L = 1u << (a.n - round);
      const T* wl = nullptr;
      const T* wr = nullptr;
      uint32_t nxr_log = 0;
      if constexpr (EQ) select_eq<FC>(round, a.n, a.eq_l_fam, a.eq_r_fam, &a.ctl->one, wl, wr, nxr_log);
      if (round == 0) {
        pass_accumulate<FS, FC, K, EQ, false>(a.in, a.in_stride, a.work, a.work_stride, L, cm, wl,
                                              wr, nxr_log, tid, THREADS, acc);
      } else if (round == 1) {
        pass_accumulate<FS, FC, K, EQ, true>(a.in, a.in_stride, a.work, a.work_stride, L, cm, wl,
                                             wr, nxr_log, tid, THREADS, acc);
      } else {
        pass_accumulate<FC, FC, K, EQ, true>(a.work, a.work_stride, a.work, a.work_stride, L, cm,
                                             wl, wr, nxr_log, tid, THREADS, acc);
      }
    } else {
      if (from_regs) {
        // T'[j], T'[j + L/2] from lane j; T'[j + P], T'[j + P + L/2] from lane j + P.
        const int src = (int)((lane + P) & 63u);
#pragma unroll
        for (int q = 0; q < K; ++q) {
          ps.s[q][0] = av[q];
          ps.s[q][1] = shfl_el<FC>(av[q], src);
          ps.s[q][2] = bv[q];
          ps.s[q][3] = shfl_el<FC>(bv[q], src);
        }
      } else if (!have_src && tid < P) {
        pair_load<FS, FC, K, EQ>(a, round, tid, ps);
      }
      if (tid < P) pair_fold_eval<FS, FC, K, EQ>(a, round, tid, ps, cm, av, bv, acc);
    }
    if (tid == 0) trace_mark(a.trace, round, 2);
    if (wave_mode) {
#pragma unroll
      for (int t = 0; t < NA; ++t) {
        T v = acc[t];
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) v = FC::add(v, shfl_xor_el<FC>(v, m));
        acc[t] = v;
      }
      if (lane == 0) {
#pragma unroll
        for (int t = 0; t < NA; ++t) sh.out[t] = acc[t];
#pragma unroll
        for (int q = 0; q < K; ++q) {
          sh.lastv[q][0] = av[q];
          sh.lastv[q][1] = bv[q];
        }
      }
      wave_sync();
    } else {
      // The folded entries are complete before the reduction's barriers.
      __builtin_amdgcn_s_waitcnt(0);
      block_reduce<FC, NA, THREADS>(acc, sh.lds, sh.out);
    }
    if (tid == 0) trace_mark(a.trace, round, 3);
    // Next round's inputs: a wave-mode round shuffles its sources out of this
    // round's pairs and needs only its weights; a one-pair round after a
    // block round loads its sources now (the tail kernel) or at its start.
    have_src = false;
    from_regs = wave_mode;
    if (round + 1 < last) {
      const uint32_t Pn = P >> 1;
      const bool next_stride = round + 1 < 2 || Pn > (uint32_t)THREADS;
      if (!next_stride && tid < Pn) {
        if (wave_mode) {
          pair_weight_load<FS, FC, K, EQ>(a, round + 1, tid, ps);
        } else if (HOP == HOP_INLINE) {
          pair_load<FS, FC, K, EQ>(a, round + 1, tid, ps);
          have_src = true;
        }
      }
    }
    if (tid < 64) {
      // The final fold reads the last pair from LDS in wave mode.
      const T* last_src = wave_mode ? &sh.lastv[0][0] : a.work;
      const size_t last_stride = wave_mode ? 2 : round == 0 ? a.in_stride : a.work_stride;
      finish_round<FS, FC, K, EQ, HOP>(a, round, round + 1 == last, sh.out, sh.sm,
                                              last_src, a.in, last_stride,
                                              !wave_mode && round == 0);
    }
    if (!wave_mode) __syncthreads();
  }
  __syncthreads();
  if (tid < 64) fin_store(sh.sm, a.ctl);
}

// Single-block kernel for the tail rounds of one proof.
template <class FS, class FC, int K, bool EQ>
__global__ __launch_bounds__(TAIL_THREADS) void tail_kernel(StageArgs<FS, FC> a, uint32_t first,
                                                            uint32_t last) {
  __shared__ TailLds<FC, K + 1, TAIL_THREADS> sh;
  tail_rounds<FS, FC, K, EQ, TAIL_THREADS>(a, first, last, sh);
}

// Slot rows of the look-ahead coefficient sums.
constexpr int LA_NSLOT = 8;     // slot rows per parity
constexpr int LA_SLOTW = 128;   // 64-bit words per row: round sums, then coefficients
constexpr size_t LA_SLOT_WORDS = (size_t)2 * LA_NSLOT * LA_SLOTW;

// Look-ahead rows of the sums of round x (parity x & 1): after the pass and
// chained kernels' slot rows.
__device__ __forceinline__ uint64_t* la_rows(uint64_t* slots, uint32_t x) {
  return slots + (size_t)NSLOT * SLOTW + (size_t)(x & 1u) * LA_NSLOT * LA_SLOTW;
}

// ---------------------------------------------------- look-ahead tail rounds
// The tail rounds [first, n) (first >= 2, at least two rounds) of one proof
// in one block, with each round's folding and evaluation run while the
// finisher wave hashes the previous round's message.
//
// Round k+1's tables are round k's folded at the challenge r_k, so each of
// their entries is a + r_k b with a, b entries of round k's tables, and the
// node sums of round k+1 are polynomials of degree K in r_k. Period P_k folds
// round k-1's tables at r_{k-1} into round k's (in place in `work`) and sums,
// per node, the coefficients C_{k+1} of that polynomial over round k+1's
// pairs; it needs only r_{k-1}, so it runs while wave 0 turns round k's
// message (from C_k and r_{k-1}) into r_k. The first period, which starts with
// the kernel, runs in two parts: every wave folds and sums round `first`'s
// node values directly, then, while the finisher hashes that message, the
// worker waves sum C_{first+1} from the folded tables. The last period
// (k = n-1) only folds, for the final fold. Thread g of a period takes round
// k+1's pair g: eight entries per table of round k-1's tables, at g + m L/8
// (L their length, m = 0..7), give the four entries of round k's at
// g + m L/8 (m = 0..3), which hold round k's pairs g and g + L/8 and, folded
// at r_k, round k+1's pair g.
//
// Wave 0 is the finisher. The first period runs on every wave; later ones on
// waves 1-3 and 5-7, off the SIMD of wave 0 (waves are dealt to the four SIMDs
// in turn), whose hash is bound by its own instruction issue. Each wave of a
// period reduces its sums into an LDS row of the period's parity and arrives
// on an LDS counter; the finisher waits for every arrival of period k before
// publishing r_k and the start of period k+1 (period k+1 reads round k's
// tables, complete only then). Every wait is bounded; one that reaches its
// limit marks the proof faulted.
constexpr int LA_THREADS = 512;
constexpr uint32_t LA_FIRST_WAVES = LA_THREADS / 64;  // waves of the first period
constexpr uint32_t LA_WAVES = 5;                      // waves of the later periods
constexpr uint32_t LA_SPIN_LIMIT = 1u << 20;

template <class FC, int K>
struct TailLaLds {
  static constexpr int NA = K + 1;
  static constexpr int NP = NA * (K + 2);  // per wave: round sums (NA), then coefficients
  FinishLds<FC> sm;
  typename FC::T part[2][LA_FIRST_WAVES][NP];
  typename FC::T coef[NA * NA];
  uint64_t gsum[NA * NA * 4];  // round first's coefficient words from the slot rows (cont)
  typename FC::T out[NA];
  typename FC::T pub_r;  // r_{k-1} for period k
  uint32_t go;           // periods allowed to start: every k < go
  uint32_t arrived;      // wave arrivals over all periods so far
  uint32_t fault;
  // The helper wave's side (rounds after the first): round k's message as
  // polynomials in r_{k-1} (by parity), the coefficients it sums them from,
  // the messages and challenges by parity for its claim and eq bookkeeping.
  typename FC::T mp[2][MAXPTS][MAXPTS + 1];
  typename FC::T pnat[MAXPTS][MAXPTS + 1];
  typename FC::T hcoef[NA * NA];
  typename FC::T msgv[2][MAXPTS];
  typename FC::T rpub[2];
  uint32_t mp_ready;  // rounds whose message polynomials are in mp: every k <= mp_ready
  uint32_t booked;    // rounds whose eq and claim bookkeeping is done: every k < booked
};

// Round k's message as polynomials in x = r_{k-1} (the tail's helper wave,
// all lanes): node t's sum is V_t(x) = sum_i C[t][i] x^i (C the coefficients
// of round k's node sums, t <= K; the extension node's V is the extension of
// those), and with eq the message value is l_{k-1}(x) E L_t V_t(x), where E
// is the eq product before round k-1, l_{k-1}(x) = B0 + x A0 its line and
// L_t = B1 + node_t A1 round k's line at node t. mp[w][j] is the coefficient
// of x^j of wire element w (row w of the wire map applied to the nodes').
template <class FC, int K, bool EQ>
__device__ __forceinline__ void tail_msg_poly(const FinishLds<FC>& sm, const typename FC::T* C,
                                              typename FC::T E, typename FC::T A0,
                                              typename FC::T B0, typename FC::T A1,
                                              typename FC::T B1,
                                              typename FC::T (*mp)[MAXPTS + 1],
                                              typename FC::T (*pn)[MAXPTS + 1]) {
  using T = typename FC::T;
  constexpr int NA = K + 1;
  constexpr int NPTS = K + 1 + (EQ ? 1 : 0);
  constexpr int D = EQ ? K + 1 : K;  // degree in x
  const uint32_t lane = threadIdx.x & 63u;
  const bool values = sm.wire == WIRE_VALUES;
  T (*dst)[MAXPTS + 1] = values ? mp : pn;
  auto vt = [&](uint32_t t, int i) -> T {
    if (i < 0 || i > K) return FC::zero();
    if (t <= (uint32_t)K) return C[t * NA + i];
    T v = FC::zero();
#pragma unroll
    for (int u = 0; u < NA; ++u) v = FC::add(v, FC::mul(sm.ext_coef[u], C[u * NA + i]));
    return v;
  };
  for (uint32_t idx = lane; idx < (uint32_t)(NPTS * (D + 1)); idx += 64u) {
    const uint32_t t = idx / (D + 1);
    const int j = (int)(idx % (D + 1));
    T v;
    if constexpr (EQ) {
      const T Lt = FC::add(B1, FC::mul(sm.nodes[t], A1));
      const T e0 = FC::mul(E, B0), e1 = FC::mul(E, A0);
      v = FC::mul(Lt, FC::add(FC::mul(e0, vt(t, j)), FC::mul(e1, vt(t, j - 1))));
    } else {
      v = vt(t, j);
    }
    dst[t][j] = v;
  }
  wave_sync();
  if (!values) {
    for (uint32_t idx = lane; idx < sm.mlen * (uint32_t)(D + 1); idx += 64u) {
      const uint32_t w = idx / (D + 1), j = idx % (D + 1);
      T v = FC::zero();
#pragma unroll
      for (int t = 0; t < NPTS; ++t) v = FC::add(v, FC::mul(sm.wmat[w][t], pn[t][j]));
      mp[w][j] = v;
    }
    wave_sync();
  }
}

// Spin on an LDS word until it reaches `target` (relaxed polls, then an
// acquire fence for the work-group). False when the limit is reached.
__device__ __forceinline__ bool la_wait(uint32_t* p, uint32_t target) {
  for (uint32_t i = 0; i < LA_SPIN_LIMIT; ++i) {
    if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= target) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      return true;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return false;
}

// la_wait that also gives up (false) once `fault` is set.
__device__ __forceinline__ bool la_wait_f(uint32_t* p, uint32_t target, uint32_t* fault) {
  for (uint32_t i = 0; i < LA_SPIN_LIMIT; ++i) {
    if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= target) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      return true;
    }
    if (__hip_atomic_load(fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) return false;
    __builtin_amdgcn_s_sleep(1);
  }
  return false;
}

// Eq weight of round `round`'s pair j.
template <class FC>
__device__ __forceinline__ typename FC::T la_weight(const typename FC::T* wl,
                                                    const typename FC::T* wr, uint32_t nxr_log,
                                                    uint32_t j) {
  if (wl) return FC::mul(wl[j >> nxr_log], wr[j & ((1u << nxr_log) - 1u)]);
  return wr[j];
}

// Coefficients of prod_q (A_q + r B_q) in r, degree K.
template <class FC, int K>
__device__ __forceinline__ void la_poly(const typename FC::T (&A)[K], const typename FC::T (&B)[K],
                                        typename FC::T (&c)[K + 1]) {
  using T = typename FC::T;
  if constexpr (K == 2) {
    c[0] = FC::mul(A[0], A[1]);
    c[2] = FC::mul(B[0], B[1]);
    const T m = FC::mul(FC::add(A[0], B[0]), FC::add(A[1], B[1]));
    c[1] = FC::sub(FC::sub(m, c[0]), c[2]);
  } else {
    c[0] = A[0];
    c[1] = B[0];
#pragma unroll
    for (int i = 2; i <= K; ++i) c[i] = FC::zero();
#pragma unroll
    for (int q = 1; q < K; ++q) {
      // (c0 + c1 r + ... + c_q r^q)(A_q + r B_q)
#pragma unroll
      for (int i = q + 1; i >= 1; --i) c[i] = FC::add(FC::mul(c[i], A[q]), FC::mul(c[i - 1], B[q]));
      c[0] = FC::mul(c[0], A[q]);
    }
  }
}

// Parts of a period: fold round k-1's tables at cm.r into round k's and sum
// round k's node values into d (LA_FOLD_D); sum the coefficients of round
// k+1's node sums into c[t * (K+1) + i] from round k's tables (LA_C); both
// fold and coefficients (LA_FOLD_C); the fold alone (LA_FOLD, a thread per
// group of four entries, or per entry below four); fold, round sums and
// coefficients (LA_FOLD_DC).
constexpr int LA_FOLD_D = 0, LA_C = 1, LA_FOLD_C = 2, LA_FOLD = 3, LA_FOLD_DC = 4;

// Period P_k (part MODE) for thread wt of nt.
template <class FS, class FC, int K, bool EQ, int MODE>
__device__ __forceinline__ void la_period(const StageArgs<FS, FC>& a, uint32_t k, uint32_t wt,
                                          uint32_t nt, const ChalMul<FC>& cm,
                                          typename FC::T (&d)[K + 1],
                                          typename FC::T (&c)[(K + 1) * (K + 1)]) {
  using T = typename FC::T;
  const uint32_t n = a.n;
  if constexpr (MODE == LA_FOLD) {
    const uint32_t L = 1u << (n - k), per = L >= 4u ? 4u : 1u, ng = L / per;
    for (uint32_t g = wt; g < ng; g += nt) {
#pragma unroll
      for (int q = 0; q < K; ++q) {
        T* s = a.work + (size_t)q * a.work_stride;
        for (uint32_t m = 0; m < per; ++m) {
          const uint32_t y = g + m * ng;
          const T lo = s[y], hi = s[y + L];
          s[y] = FC::add(lo, cm.mul(FC::sub(hi, lo)));
        }
      }
    }
    return;
  }
  const uint32_t q8 = 1u << (n - k - 2);  // L/8 = round k+1's pairs
  const T* wl_d = nullptr;
  const T* wr_d = nullptr;
  const T* wl_c = nullptr;
  const T* wr_c = nullptr;
  uint32_t xd = 0, xc = 0;
  if constexpr (EQ) {
    select_eq<FC>(k, n, a.eq_l_fam, a.eq_r_fam, &a.ctl->one, wl_d, wr_d, xd);
    select_eq<FC>(k + 1, n, a.eq_l_fam, a.eq_r_fam, &a.ctl->one, wl_c, wr_c, xc);
  }
  for (uint32_t g = wt; g < q8; g += nt) {
    T e[K][4];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      T* s = a.work + (size_t)q * a.work_stride;
      if constexpr (MODE == LA_C) {
#pragma unroll
        for (int m = 0; m < 4; ++m) e[q][m] = s[g + (size_t)m * q8];
      } else {
        T x[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) x[m] = s[g + (size_t)m * q8];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          e[q][m] = FC::add(x[m], cm.mul(FC::sub(x[m + 4], x[m])));
          s[g + (size_t)m * q8] = e[q][m];
        }
      }
    }
    if constexpr (MODE == LA_FOLD_D || MODE == LA_FOLD_DC) {
      // Round k's pairs g = (e0, e2) and g + L/8 = (e1, e3).
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        T lo[K], df[K];
#pragma unroll
        for (int q = 0; q < K; ++q) {
          lo[q] = e[q][h];
          df[q] = FC::sub(e[q][h + 2], e[q][h]);
        }
        if constexpr (EQ) {
          const T w = la_weight<FC>(wl_d, wr_d, xd, g + (uint32_t)h * q8);
          const T hi0 = FC::add(lo[0], df[0]);
          lo[0] = FC::mul(w, lo[0]);
          df[0] = FC::sub(FC::mul(w, hi0), lo[0]);
        }
#pragma unroll
        for (int t = 0; t <= K; ++t) {
          T prod = lo[0];
#pragma unroll
          for (int q = 1; q < K; ++q) prod = FC::mul(prod, lo[q]);
          d[t] = FC::add(d[t], prod);
          if (t < K) {
#pragma unroll
            for (int q = 0; q < K; ++q) lo[q] = FC::add(lo[q], df[q]);
          }
        }
      }
      if constexpr (MODE == LA_FOLD_D) continue;
    }
    // Round k+1's pair g: lo = a0 + r b0, hi = a1 + r b1 per table.
    T A[K], B[K], dA[K], dB[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      A[q] = e[q][0];
      B[q] = FC::sub(e[q][2], e[q][0]);
      dA[q] = FC::sub(e[q][1], e[q][0]);
      dB[q] = FC::sub(FC::sub(e[q][3], e[q][1]), B[q]);
    }
    if constexpr (EQ) {
      const T w = la_weight<FC>(wl_c, wr_c, xc, g);
      A[0] = FC::mul(w, A[0]);
      B[0] = FC::mul(w, B[0]);
      dA[0] = FC::mul(w, dA[0]);
      dB[0] = FC::mul(w, dB[0]);
    }
#pragma unroll
    for (int t = 0; t <= K; ++t) {
      T p[K + 1];
      la_poly<FC, K>(A, B, p);
#pragma unroll
      for (int i = 0; i <= K; ++i) c[t * (K + 1) + i] = FC::add(c[t * (K + 1) + i], p[i]);
      if (t < K) {
#pragma unroll
        for (int q = 0; q < K; ++q) {
          A[q] = FC::add(A[q], dA[q]);
          B[q] = FC::add(B[q], dB[q]);
        }
      }
    }
  }
}

// Sum a period's values over the wave into row `row` (lane 0 writes) and
// arrive. A wave with no pairs (wave_base >= pairs) writes zeros.
template <class FC, int NV>
__device__ __forceinline__ void la_arrive(typename FC::T (&v)[NV], bool any,
                                          typename FC::T* row, uint32_t* arrived) {
  const uint32_t lane = threadIdx.x & 63;
  if (any) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      typename FC::T x = v[i];
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) x = FC::add(x, shfl_xor_el<FC>(x, m));
      v[i] = x;
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) row[i] = any ? v[i] : FC::zero();
  }
  // The row and this wave's table entries are written before the arrival.
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  if (lane == 0) __hip_atomic_fetch_add(arrived, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// With `cont`, the previous launch (chain_la_kernel) left round first's
// coefficients in the look-ahead slot rows: the first message is hashed at
// once and the first period is an ordinary one.
template <class FS, class FC, int K, bool EQ>
__global__ __launch_bounds__(LA_THREADS) void tail_la_kernel(StageArgs<FS, FC> a,
                                                             uint32_t first, uint32_t cont) {
  using T = typename FC::T;
  constexpr int NA = K + 1, NC = NA * NA;
  __shared__ TailLaLds<FC, K> sh;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t n = a.n;
  if (tid < 64) {
    fin_load(sh.sm, a.ctl);
    fin_eq_table<FS, FC, EQ>(sh.sm, a);
  }
  if (tid == 0) {
    sh.go = first + 1;
    sh.arrived = 0;
    sh.fault = 0;
    sh.pub_r = a.ctl->r;
    sh.mp_ready = first;
    sh.booked = 0;
  }
  __syncthreads();
  // Wave index among the later periods' waves (1-3, 5, 6 -> 0..4), or -1.
  // Waves w and w + 4 share a SIMD: the finisher (wave 0) has its SIMD to
  // itself, the periods take waves 1-3, 5, 6 and the helper wave 7.
  constexpr uint32_t HELPER_WAVE = 7;
  const int widx = wave == 0 || wave == 4 || wave == HELPER_WAVE
                       ? -1
                       : (int)(wave < 4 ? wave - 1 : wave - 2);
  const uint32_t q8f = 1u << (n - first - 2);  // pairs of round first+1
  // Every wave: the first period's fold and round first's node sums.
  if (!cont) {
    ChalMul<FC> cm;
    cm.init(sh.pub_r);
    T d[NA], c[NC];
#pragma unroll
    for (int t = 0; t < NA; ++t) d[t] = FC::zero();
    la_period<FS, FC, K, EQ, LA_FOLD_D>(a, first, tid, LA_THREADS, cm, d, c);
    la_arrive<FC, NA>(d, wave * 64u < q8f, sh.part[first & 1][wave], &sh.arrived);
  }
  if (widx >= 0) {
    // Round first+1's coefficients, once every wave has folded.
    if (!cont && !la_wait(&sh.arrived, LA_FIRST_WAVES)) {
      sh.fault = 1;
      return;
    }
    if (!cont) {
      ChalMul<FC> cm;
      T d[NA], c[NC];
#pragma unroll
      for (int i = 0; i < NC; ++i) c[i] = FC::zero();
      const uint32_t wt = (uint32_t)widx * 64u + lane;
      la_period<FS, FC, K, EQ, LA_C>(a, first, wt, LA_WAVES * 64u, cm, d, c);
      la_arrive<FC, NC>(c, (uint32_t)widx * 64u < q8f, sh.part[first & 1][widx] + NA,
                        &sh.arrived);
    }
    // Later periods (with `cont`, from the first).
    for (uint32_t k = cont ? first : first + 1; k < n; ++k) {
      if (!la_wait(&sh.go, k + 1)) {
        sh.fault = 1;
        return;
      }
      if (__hip_atomic_load(&sh.fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) return;
      ChalMul<FC> cm;
      cm.init(sh.pub_r);
      T d[NA], c[NC];
#pragma unroll
      for (int t = 0; t < NA; ++t) d[t] = FC::zero();
#pragma unroll
      for (int i = 0; i < NC; ++i) c[i] = FC::zero();
      const uint32_t wt = (uint32_t)widx * 64u + lane;
      if (k + 1 < n)
        la_period<FS, FC, K, EQ, LA_FOLD_C>(a, k, wt, LA_WAVES * 64u, cm, d, c);
      else
        la_period<FS, FC, K, EQ, LA_FOLD>(a, k, wt, LA_WAVES * 64u, cm, d, c);
      const uint32_t q8 = k + 2 <= n ? 1u << (n - k - 2) : 0u;
      la_arrive<FC, NC>(c, k + 1 < n && (uint32_t)widx * 64u < q8,
                        sh.part[k & 1][widx] + NA, &sh.arrived);
    }
    return;
  }
  constexpr int NPTS = K + 1 + (EQ ? 1 : 0);
  constexpr int D = EQ ? K + 1 : K;  // degree of a message polynomial
  const uint32_t base_t = cont ? 0u : LA_FIRST_WAVES;
  if (wave == HELPER_WAVE) {
    // The helper: while the finisher hashes round k-1, form round k's
    // message as polynomials in r_{k-1} from period k-1's coefficients, so
    // that the finisher only evaluates them at r_{k-1}; once r_{k-1} is out,
    // the eq product and (when needed) the claim of round k-1.
    if (first + 1 >= n) return;
    T E = EQ ? a.ctl->eq_prev : FC::one();  // eq product before round k-1
    T claim = FC::zero();
    for (uint32_t k = first + 1; k <= n; ++k) {
      if (k < n) {
        if (!la_wait_f(&sh.arrived, base_t + (k - first) * LA_WAVES, &sh.fault)) break;
        if (lane < (uint32_t)NC) {
          T v = FC::zero();
          for (uint32_t w = 0; w < LA_WAVES; ++w) v = FC::add(v, sh.part[(k - 1) & 1][w][NA + lane]);
          sh.hcoef[lane] = v;
        }
        wave_sync();
        T A0 = FC::zero(), B0 = FC::zero(), A1 = FC::zero(), B1 = FC::zero();
        if constexpr (EQ) {
          A0 = sh.sm.eqab[2 * (k - 1)];
          B0 = sh.sm.eqab[2 * (k - 1) + 1];
          A1 = sh.sm.eqab[2 * k];
          B1 = sh.sm.eqab[2 * k + 1];
        }
        tail_msg_poly<FC, K, EQ>(sh.sm, sh.hcoef, E, A0, B0, A1, B1, sh.mp[k & 1], sh.pnat);
        if (lane == 0) {
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
          __hip_atomic_store(&sh.mp_ready, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
      if (!la_wait_f(&sh.go, k + 1, &sh.fault)) break;
      const T r = sh.rpub[(k - 1) & 1];
      if (k - 1 == first) {
        claim = sh.sm.claim;  // the finisher's own reduction of round `first`
      } else if (k == n || sh.sm.claim_every) {
        claim = round_claim<FC, NPTS>(sh.sm, sh.msgv[(k - 1) & 1], r, claim);
      }
      if constexpr (EQ)
        E = FC::mul(E, FC::add(sh.sm.eqab[2 * (k - 1) + 1], FC::mul(r, sh.sm.eqab[2 * (k - 1)])));
      if (k == n && lane == 0) {
        sh.sm.eq_prev = E;
        sh.sm.claim = claim;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __hip_atomic_store(&sh.booked, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      wave_sync();
    }
    if (__hip_atomic_load(&sh.booked, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < n && lane == 0)
      __hip_atomic_store(&sh.fault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    return;
  }
  if (wave != 0) return;
  // The finisher.
  uint32_t target = base_t;
  bool ok = true;
  for (uint32_t k = first; k < n && ok; ++k) {
    if (lane == 0) trace_mark(a.trace, k, 1);
    fin_eq_lds<FC, EQ>(sh.sm, k);
    if (k == first && cont) {
      // Round first's node sums: the previous launch's coefficients at
      // r_{first-1} (complete at the launch boundary: plain loads, all rows
      // at once).
      constexpr int W = SlotOps<FC>::W;
      uint64_t* rows = la_rows(a.slots, first) + NA * W;
      for (uint32_t w = lane; w < (uint32_t)(NC * W); w += 64u) {
        uint64_t v[LA_NSLOT];
#pragma unroll
        for (int row = 0; row < LA_NSLOT; ++row) v[row] = rows[(size_t)row * LA_SLOTW + w];
        uint64_t s = 0;
#pragma unroll
        for (int row = 0; row < LA_NSLOT; ++row) {
          s += v[row];
          rows[(size_t)row * LA_SLOTW + w] = 0;
        }
        sh.gsum[w] = s;
      }
      wave_sync();
      if (lane < (uint32_t)NC) sh.coef[lane] = SlotOps<FC>::join(&sh.gsum[lane * W]);
      wave_sync();
      if (lane < (uint32_t)NA) {
        const T r = sh.sm.r;
        T v = sh.coef[lane * NA + K];
#pragma unroll
        for (int i = K - 1; i >= 0; --i) v = FC::add(sh.coef[lane * NA + i], FC::mul(r, v));
        sh.out[lane] = v;
      }
    } else if (k == first) {
      ok = la_wait(&sh.arrived, target);
      if (!ok) break;
      // Round first's node sums, from every wave of the first period.
      if (lane < (uint32_t)NA) {
        T s = FC::zero();
        for (uint32_t w = 0; w < LA_FIRST_WAVES; ++w) s = FC::add(s, sh.part[first & 1][w][lane]);
        sh.out[lane] = s;
      }
    } else {
      // Round k's message: the helper's polynomials at r_{k-1}; the hop;
      // the challenge (the helper keeps the eq product and the claim).
      ok = la_wait_f(&sh.mp_ready, k, &sh.fault);
      if (!ok) break;
      const uint32_t par = k & 1u;
      const uint32_t mlen = sh.sm.mlen;
      if (lane < mlen) {
        const T r = sh.sm.r;
        T v = sh.mp[par][lane][D];
#pragma unroll
        for (int j = D - 1; j >= 0; --j) v = FC::add(FC::mul(v, r), sh.mp[par][lane][j]);
        a.msgs[(size_t)k * NPTS + lane] = v;
        sh.sm.msg[lane] = v;
        sh.msgv[par][lane] = v;
      }
      wave_sync();
      if (lane == 0) trace_mark(a.trace, k, 4);
      fs_hop_body(sh.sm, reinterpret_cast<const uint8_t*>(sh.sm.msg), mlen * (uint32_t)sizeof(T),
                  a.ctl->item_bytes, a.ctl->limbs, a.ctl->tr_kind);
      if (lane == 0) trace_mark(a.trace, k, 5);
      if (lane < (uint32_t)sizeof(T)) a.point[(size_t)k * sizeof(T) + lane] = sh.sm.out[lane];
      const T r = el_from_words<FC>(reinterpret_cast<const uint32_t*>(sh.sm.out));
      if (lane == 0) {
        sh.sm.r = r;
        trace_mark(a.trace, k, 6);
      }
      wave_sync();
    }
    if (k == first) {
      wave_sync();
      StageArgs<FS, FC> af = a;
      af.final_out = nullptr;  // the final fold waits for the last period
      finish_round<FS, FC, K, EQ, HOP_INLINE>(af, k, k + 1 == n, sh.out, sh.sm, a.work, a.in,
                                              a.work_stride, false);
    }
    // Period k is complete: publish r_k (for the helper too) and start
    // period k+1.
    target += LA_WAVES;
    ok = la_wait(&sh.arrived, target);
    if (!ok) break;
    if (lane == 0) {
      sh.pub_r = sh.sm.r;
      sh.rpub[k & 1u] = sh.sm.r;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __hip_atomic_store(&sh.go, k + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    wave_sync();
  }
  if (ok && first + 1 < n) ok = la_wait_f(&sh.booked, n, &sh.fault);
  if (!ok) {
    if (lane == 0) {
      __hip_atomic_store(&sh.fault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_store(&a.ctl->fault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return;
  }
  // The final fold of round n-1's entries (written by the last period).
  if (a.final_out && lane < (uint32_t)K) {
    const T* s = a.work + (size_t)lane * a.work_stride;
    const T lo = s[0], hi = s[1];
    a.final_out[lane] = FC::add(lo, FC::mul(sh.sm.r, FC::sub(hi, lo)));
  }
  if (EQ && a.final_out && lane == 0) a.final_out[K] = sh.sm.eq_prev;
  fin_store(sh.sm, a.ctl);
}

// ------------------------------------------------ look-ahead chained rounds
// Rounds [first, last) (first >= 2, at least two rounds) of one proof in one
// persistent launch, with the periods of the look-ahead tail spread over
// blocks 1..G (all resident) and block 0 as the finisher: period P_k folds
// round k-1's tables into round k's and sums the coefficients C_{k+1} of
// round k+1's node sums (polynomials in r_k) while block 0 hashes round k's
// message, which it forms from C_k and r_{k-1}. The first period folds and
// sums round `first`'s node values directly, then sums C_{first+1} from the
// folded tables while that message is hashed; with `cont` the previous
// launch left C_first and the first period is an ordinary one. The period
// of round last-1 also sums round last's coefficients, for the tail launch
// that follows; when last = n it only folds, for block 0's final fold.
// Cross-block sums go through two parities of slot rows (the sums of round x
// land in parity x & 1, and block 0 reads and clears round k's before it
// lets period k+1, which writes round k+2's, start); every block that works
// in a period arrives at the control block's counter (release), and block 0
// publishes r_k (`done`, release) once every block has arrived for period k,
// whose tables period k+1 reads. The claim is reduced at round
// claim_last - 1 only. Waits are bounded as in chain_kernel.

// Blocks of a period with `units` thread units of work: one per PASS_THREADS,
// at least one, at most grid.
__device__ __forceinline__ uint32_t la_blocks(uint32_t units, uint32_t grid) {
  const uint32_t b = units / PASS_THREADS;
  return b < 1u ? 1u : b < grid ? b : grid;
}

// Thread units of period k (n rounds): round k+1's pairs, or for the
// fold-only period of round n-1 its entries.
__device__ __forceinline__ uint32_t la_units(uint32_t n, uint32_t k) {
  if (k + 1 < n) return 1u << (n - k - 2);
  const uint32_t L = 1u << (n - k);
  return L >= 4u ? L / 4u : L;
}

template <class FC, int NV>
__device__ __forceinline__ void la_slot_add(uint64_t* rows, uint32_t blk, uint32_t word_off,
                                            const typename FC::T* vals) {
  constexpr int W = SlotOps<FC>::W;
  static_assert(NV * W <= LA_SLOTW, "look-ahead slot row width");
  for (uint32_t i = threadIdx.x; i < (uint32_t)(NV * W); i += PASS_THREADS) {
    uint64_t* slot = rows + (size_t)(blk % LA_NSLOT) * LA_SLOTW + word_off + i;
    __hip_atomic_fetch_add(slot, (uint64_t)SlotOps<FC>::word(vals[i / W], (int)(i % W)),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Sum (and clear) words [word_off, word_off + NV W) of the rows into out[0..NV).
// Whole block; ends with a barrier.
template <class FC, int NV>
__device__ __forceinline__ void la_slot_sum(uint64_t* rows, uint32_t word_off,
                                            uint64_t (*sum64)[LA_SLOTW], typename FC::T* out) {
  constexpr int W = SlotOps<FC>::W;
  constexpr int G = PASS_THREADS / LA_SLOTW;
  static_assert(LA_NSLOT % G == 0, "look-ahead slot layout");
  const uint32_t w = threadIdx.x % LA_SLOTW, g = threadIdx.x / LA_SLOTW;
  uint64_t s = 0;
  if (w < (uint32_t)(NV * W)) {
#pragma unroll
    for (int i = 0; i < LA_NSLOT / G; ++i) {
      uint64_t* p = rows + (size_t)(g + i * G) * LA_SLOTW + word_off + w;
      s += __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(p, (uint64_t)0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  sum64[g][w] = s;
  __syncthreads();
  if (threadIdx.x < (uint32_t)NV) {
    uint64_t tot[W];
#pragma unroll
    for (int c = 0; c < W; ++c) {
      uint64_t v = 0;
#pragma unroll
      for (int q = 0; q < G; ++q) v += sum64[q][threadIdx.x * W + c];
      tot[c] = v;
    }
    out[threadIdx.x] = SlotOps<FC>::join(tot);
  }
  __syncthreads();
}

// la_slot_sum for one wave (all lanes): words [word_off, word_off + NV W)
// of the rows into out[0..NV), clearing them; `scratch` (LDS) holds NV W
// words.
template <class FC, int NV>
__device__ __forceinline__ void la_slot_sum_wave(uint64_t* rows, uint32_t word_off,
                                                 uint64_t* scratch, typename FC::T* out) {
  constexpr int W = SlotOps<FC>::W;
  const uint32_t lane = threadIdx.x & 63u;
  for (uint32_t w = lane; w < (uint32_t)(NV * W); w += 64u) {
    uint64_t v[LA_NSLOT];
#pragma unroll
    for (int i = 0; i < LA_NSLOT; ++i)
      v[i] = __hip_atomic_load(rows + (size_t)i * LA_SLOTW + word_off + w, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    uint64_t t = 0;
#pragma unroll
    for (int i = 0; i < LA_NSLOT; ++i) {
      t += v[i];
      __hip_atomic_store(rows + (size_t)i * LA_SLOTW + word_off + w, (uint64_t)0,
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    scratch[w] = t;
  }
  wave_sync();
  if (lane < (uint32_t)NV) out[lane] = SlotOps<FC>::join(&scratch[lane * W]);
  wave_sync();
}

template <class FC, int K>
struct ChainLaLds {
  static constexpr int NA = K + 1, NC = NA * NA;
  FinishLds<FC> sm;
  typename FC::T lds[(PASS_THREADS / 64) * NC];
  typename FC::T out[NC];
  uint64_t sum64[PASS_THREADS / LA_SLOTW][LA_SLOTW];
  typename FC::T r;
  uint32_t ok;
  // Block 0: the finisher's slot words, and the helper wave's side as in the
  // look-ahead tail.
  uint64_t fsum[NC * 4];
  uint64_t hsum[NC * 4];
  typename FC::T mp[2][MAXPTS][MAXPTS + 1];
  typename FC::T pnat[MAXPTS][MAXPTS + 1];
  typename FC::T hcoef[NC];
  typename FC::T msgv[2][MAXPTS];
  typename FC::T rpub[2];
  uint32_t mp_ready;  // rounds whose message polynomials are in mp: every k <= mp_ready
  uint32_t pdone;     // periods complete: every k < pdone
  uint32_t go_h;      // rounds whose challenge is in rpub: every k < go_h
  uint32_t booked;    // rounds whose eq and claim bookkeeping is done: every k < booked
  uint32_t hfault;
};

template <class FT, class FC, int K, bool EQ>
__global__ __launch_bounds__(PASS_THREADS) void chain_la_kernel(
    StageArgs<FT, FC> a, uint32_t first, uint32_t last, uint32_t claim_last, uint32_t cont) {
  using T = typename FC::T;
  constexpr int NA = K + 1, NC = NA * NA;
  constexpr int W = SlotOps<FC>::W;
  __shared__ ChainLaLds<FC, K> sh;
  const uint32_t tid = threadIdx.x, n = a.n;
  const uint32_t G = gridDim.x - 1;
  Ctl<FC>* ctl = a.ctl;
  // Blocks of the first period (the busiest): with `cont`, P_first folds and
  // sums coefficients; else its fold-and-sums part, then its coefficients.
  const uint32_t nb0 = la_blocks(cont ? la_units(n, first) : 1u << (n - first - 2), G);
  if (blockIdx.x == 0) {
    // Block 0: the finisher (wave 0) and its helper (wave 1, another SIMD).
    // The helper waits for each period's arrivals, sums the coefficients it
    // left, forms the next round's message as polynomials in the pending
    // challenge (tail_msg_poly) and keeps the eq product and the claim, so
    // that after each hop the finisher only publishes the challenge and
    // evaluates the next message.
    constexpr int NPTS = K + 1 + (EQ ? 1 : 0);
    constexpr int D = EQ ? K + 1 : K;
    const uint32_t wave = tid >> 6, lane = tid & 63u;
    if (tid < 64) {
      fin_load(sh.sm, ctl);
      fin_eq_table<FT, FC, EQ>(sh.sm, a);
    }
    if (tid == 0) {
      trace_mark(a.trace, first, 7);
      sh.mp_ready = first;
      sh.pdone = first;
      sh.go_h = first;
      sh.booked = 0;
      sh.hfault = 0;
    }
    __syncthreads();
    if (wave == 1) {
      T E = EQ ? ctl->eq_prev : FC::one();  // eq product before round k-1
      T claim = FC::zero();
      uint32_t target = cont ? 0u : nb0;
      bool ok = true;
      for (uint32_t k = first + 1; k <= last && ok; ++k) {
        // Period k-1 complete; round k's message polynomials from the
        // coefficients it left.
        target += k - 1 == first ? nb0 : la_blocks(la_units(n, k - 1), G);
        ok = chain_wait(&ctl->counter, target, &ctl->fault);
        if (!ok) break;
        // At once: the finisher publishes r_{k-1} (the start of period k) on it.
        if (lane == 0) {
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
          __hip_atomic_store(&sh.pdone, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        if (k < last) {
          la_slot_sum_wave<FC, NC>(la_rows(a.slots, k), NA * W, sh.hsum, sh.hcoef);
          T A0 = FC::zero(), B0 = FC::zero(), A1 = FC::zero(), B1 = FC::zero();
          if constexpr (EQ) {
            A0 = sh.sm.eqab[2 * (k - 1)];
            B0 = sh.sm.eqab[2 * (k - 1) + 1];
            A1 = sh.sm.eqab[2 * k];
            B1 = sh.sm.eqab[2 * k + 1];
          }
          tail_msg_poly<FC, K, EQ>(sh.sm, sh.hcoef, E, A0, B0, A1, B1, sh.mp[k & 1], sh.pnat);
        }
        if (k < last && lane == 0) {
          // The cleared rows are visible before the additions of period k+1
          // (which follow the finisher's publication of r_k).
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
          __hip_atomic_store(&sh.mp_ready, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        // The eq product and the claim of round k-1, once r_{k-1} is out.
        ok = la_wait_f(&sh.go_h, k, &sh.hfault);
        if (!ok) break;
        const T r = sh.rpub[(k - 1) & 1];
        if (k - 1 == first) {
          claim = sh.sm.claim;  // the finisher's own reduction of round `first`
        } else if (k == claim_last || sh.sm.claim_every) {
          claim = round_claim<FC, NPTS>(sh.sm, sh.msgv[(k - 1) & 1], r, claim);
        }
        if constexpr (EQ)
          E = FC::mul(E, FC::add(sh.sm.eqab[2 * (k - 1) + 1], FC::mul(r, sh.sm.eqab[2 * (k - 1)])));
        wave_sync();
      }
      if (lane == 0) {
        if (ok) {
          sh.sm.eq_prev = E;
          sh.sm.claim = claim;
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
          __hip_atomic_store(&sh.booked, last, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
          __hip_atomic_store(&sh.hfault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
      return;
    }
    if (wave != 0) return;
    bool ok = true;
    for (uint32_t k = first; k < last; ++k) {
      fin_eq_lds<FC, EQ>(sh.sm, k);
      if (k == first) {
        if (!cont) {
          ok = chain_wait(&ctl->counter, nb0, &ctl->fault);
          if (!ok) break;
          la_slot_sum_wave<FC, NA>(la_rows(a.slots, first), 0, sh.fsum, sh.out);
        } else {
          // Round first's node sums: the previous launch's coefficients at
          // r_{first-1}.
          la_slot_sum_wave<FC, NC>(la_rows(a.slots, first), NA * W, sh.fsum, sh.lds);
          if (lane < (uint32_t)NA) {
            const T r = sh.sm.r;
            T v = sh.lds[lane * NA + K];
#pragma unroll
            for (int i = K - 1; i >= 0; --i) v = FC::add(sh.lds[lane * NA + i], FC::mul(r, v));
            sh.out[lane] = v;
          }
          wave_sync();
        }
        if (lane == 0) trace_mark(a.trace, k, 2);
        StageArgs<FT, FC> af = a;
        af.final_out = nullptr;  // the final fold waits for the last period
        finish_round<FT, FC, K, EQ, HOP_CHAIN>(af, k, k + 1 == claim_last, sh.out, sh.sm, a.work,
                                               a.in, a.work_stride, false);
      } else {
        // Round k's message: the helper's polynomials at r_{k-1}; the hop.
        ok = la_wait_f(&sh.mp_ready, k, &sh.hfault);
        if (!ok) break;
        const uint32_t par = k & 1u;
        const uint32_t mlen = sh.sm.mlen;
        if (lane < mlen) {
          const T r = sh.sm.r;
          T v = sh.mp[par][lane][D];
#pragma unroll
          for (int j = D - 1; j >= 0; --j) v = FC::add(FC::mul(v, r), sh.mp[par][lane][j]);
          a.msgs[(size_t)k * NPTS + lane] = v;
          sh.sm.msg[lane] = v;
          sh.msgv[par][lane] = v;
        }
        wave_sync();
        if (lane == 0) trace_mark(a.trace, k, 4);
        fs_hop<HOP_CHAIN>((ZS_LDS HopLds*)static_cast<HopLds*>(&sh.sm),
                          (const ZS_LDS uint8_t*)sh.sm.msg, mlen * (uint32_t)sizeof(T),
                          ctl->item_bytes, ctl->limbs, ctl->tr_kind);
        if (lane == 0) trace_mark(a.trace, k, 5);
        if (lane < (uint32_t)sizeof(T)) a.point[(size_t)k * sizeof(T) + lane] = sh.sm.out[lane];
        const T r = el_from_words<FC>(reinterpret_cast<const uint32_t*>(sh.sm.out));
        if (lane == 0) {
          sh.sm.r = r;
          trace_mark(a.trace, k, 6);
        }
        wave_sync();
      }
      // Period k complete (the helper saw its arrivals): publish r_k, which
      // starts period k+1.
      ok = la_wait_f(&sh.pdone, k + 1, &sh.hfault);
      if (!ok) break;
      if (lane == 0) {
        if (k + 1 < last) {
          ctl->r = sh.sm.r;
          __hip_atomic_store(&ctl->done, k - first + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        sh.rpub[k & 1u] = sh.sm.r;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __hip_atomic_store(&sh.go_h, k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      wave_sync();
    }
    if (ok) ok = la_wait_f(&sh.booked, last, &sh.hfault);
    if (!ok) {
      if (lane == 0) {
        __hip_atomic_store(&sh.hfault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_store(&ctl->fault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      return;
    }
    if (last == n && a.final_out && lane < (uint32_t)K) {
      const T* s = a.work + (size_t)lane * a.work_stride;
      const T lo = s[0], hi = s[1];
      a.final_out[lane] = FC::add(lo, FC::mul(sh.sm.r, FC::sub(hi, lo)));
    }
    if (EQ && last == n && a.final_out && lane == 0) a.final_out[K] = sh.sm.eq_prev;
    fin_store(sh.sm, ctl);
    if (lane == 0) {
      __hip_atomic_store(&ctl->counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ctl->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return;
  }
  const uint32_t blk = blockIdx.x - 1;
  if (blk >= nb0) return;  // the first period has the most work
  auto arrive = [&]() {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) __hip_atomic_fetch_add(&ctl->counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  };
  if (!cont) {
    // The first period: fold and round first's node sums ...
    ChalMul<FC> cm;
    cm.init(ctl->r);
    T d[NA], c[NC];
#pragma unroll
    for (int t = 0; t < NA; ++t) d[t] = FC::zero();
    la_period<FT, FC, K, EQ, LA_FOLD_D>(a, first, blk * PASS_THREADS + tid, nb0 * PASS_THREADS,
                                        cm, d, c);
    block_reduce<FC, NA, PASS_THREADS>(d, sh.lds, sh.out);
    la_slot_add<FC, NA>(la_rows(a.slots, first), blk, 0, sh.out);
    arrive();
    // ... then, once every block has folded, round first+1's coefficients.
    if (tid == 0) sh.ok = chain_wait(&ctl->counter, nb0, &ctl->fault) ? 1u : 0u;
    __syncthreads();
    if (!sh.ok) return;
#pragma unroll
    for (int i = 0; i < NC; ++i) c[i] = FC::zero();
    la_period<FT, FC, K, EQ, LA_C>(a, first, blk * PASS_THREADS + tid, nb0 * PASS_THREADS, cm,
                                   d, c);
    block_reduce<FC, NC, PASS_THREADS>(c, sh.lds, sh.out);
    la_slot_add<FC, NC>(la_rows(a.slots, first + 1), blk, NA * W, sh.out);
    arrive();
  }
  for (uint32_t k = cont ? first : first + 1; k < last; ++k) {
    const uint32_t nb = la_blocks(la_units(n, k), G);
    if (blk >= nb) return;
    if (tid == 0) {
      const bool ok = chain_wait(&ctl->done, k - first, &ctl->fault);
      if (ok) sh.r = ctl->r;
      sh.ok = ok ? 1u : 0u;
    }
    __syncthreads();
    if (!sh.ok) return;
    ChalMul<FC> cm;
    cm.init(sh.r);
    T d[NA], c[NC];
    const uint32_t wt = blk * PASS_THREADS + tid, nt = nb * PASS_THREADS;
    if (k + 1 < n) {
#pragma unroll
      for (int i = 0; i < NC; ++i) c[i] = FC::zero();
      la_period<FT, FC, K, EQ, LA_FOLD_C>(a, k, wt, nt, cm, d, c);
      block_reduce<FC, NC, PASS_THREADS>(c, sh.lds, sh.out);
      la_slot_add<FC, NC>(la_rows(a.slots, k + 1), blk, NA * W, sh.out);
    } else {
      la_period<FT, FC, K, EQ, LA_FOLD>(a, k, wt, nt, cm, d, c);
    }
    arrive();
  }
}

// Round k >= 2 of one proof as 1 + nblk blocks, starting the look-ahead:
// blocks 1..nblk fold round k-1's tables, sum round k's node values into the
// slot rows and round k+1's coefficients into the look-ahead rows of round
// k+1, and arrive; block 0 finishes round k from the node sums as in
// pass_kernel. Round k+1 (k + 1 < n) starts a look-ahead chain with `cont`.
template <class FT, class FC, int K, bool EQ>
__global__ __launch_bounds__(PASS_THREADS) PASS_WAVES_PER_EU void pass_la_kernel(
    StageArgs<FT, FC> a, uint32_t round) {
  using T = typename FC::T;
  constexpr int NA = K + 1, NC = NA * NA;
  constexpr int NPTS = K + 1 + (EQ ? 1 : 0);
  constexpr int W = SlotOps<FC>::W;
  __shared__ PassLds<FC, NC> ps;
  RoundLds<FC, NC>& sh = ps.round;
  const uint32_t tid = threadIdx.x, nblk = gridDim.x - 1;
  Ctl<FC>* ctl = a.ctl;
  if (blockIdx.x == 0) {
    if (tid < 64) {
      fin_eq<FT, FC, EQ>(sh.sm, a, round);
      fin_load(sh.sm, ctl);
      // Warm-up hop on a copy of the stream state, as in pass_kernel.
      const uint32_t* from = reinterpret_cast<const uint32_t*>(static_cast<HopLds*>(&sh.sm));
      uint32_t* to = reinterpret_cast<uint32_t*>(&ps.warm);
      for (uint32_t i = tid; i < sizeof(HopLds) / 4; i += 64u) to[i] = from[i];
      wave_sync();
      fs_hop<HOP_PASS>((ZS_LDS HopLds*)&ps.warm, (const ZS_LDS uint8_t*)sh.sm.msg,
                       (uint32_t)(NPTS * sizeof(T)), ctl->item_bytes, ctl->limbs, ctl->tr_kind);
    }
    if (tid == 0) sh.is_last = chain_wait(&ctl->counter, nblk, &ctl->fault) ? 1u : 0u;
    __syncthreads();
    if (!sh.is_last) return;
    slot_rows_sum<FC, NA>(a.slots, sh);
    if (tid < 64) {
      StageArgs<FT, FC> af = a;
      af.final_out = nullptr;
      finish_round<FT, FC, K, EQ, HOP_PASS>(af, round, false, sh.out, sh.sm, a.work, a.in,
                                            a.work_stride, false);
      fin_store(sh.sm, ctl);
    }
    if (tid == 0) __hip_atomic_store(&ctl->counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  const uint32_t blk = blockIdx.x - 1;
  ChalMul<FC> cm;
  cm.init(ctl->r);
  T d[NA], c[NC];
#pragma unroll
  for (int t = 0; t < NA; ++t) d[t] = FC::zero();
#pragma unroll
  for (int i = 0; i < NC; ++i) c[i] = FC::zero();
  la_period<FT, FC, K, EQ, LA_FOLD_DC>(a, round, blk * PASS_THREADS + tid, nblk * PASS_THREADS,
                                       cm, d, c);
  block_reduce<FC, NA, PASS_THREADS>(d, sh.lds, sh.out);
  slot_row_add<FC, NA>(a.slots, blk, sh.out);
  block_reduce<FC, NC, PASS_THREADS>(c, sh.lds, sh.out);
  la_slot_add<FC, NC>(la_rows(a.slots, round + 1), blk, NA * W, sh.out);
  // The additions are complete before the arrival; the folded tables and the
  // coefficients reach the next launch through the kernel boundary.
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (tid == 0) __hip_atomic_fetch_add(&ctl->counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------ rounds 0-2 from the input
// The first folded tables are never written. Round 1's tables are round 0's
// folded at r_0, so round 1's node sums are polynomials of degree K in r_0
// whose coefficients are sums of products of input entries (base-field work,
// weighted by eq lazily): fused01_kernel streams the input once, sums round
// 0's node values and round 1's coefficients, and its finisher block hashes
// round 0's message, forms round 1's from the coefficients at r_0 and hashes
// it too. fold2_kernel then reads the input again and folds it twice (at r_0
// and r_1) into round 2's tables, which it writes and evaluates. Against
// writing round 1's tables in the challenge field and reading them back,
// this drops a write and a read of twice the input's size (BabyBear under
// BabyBear^4) and one launch.
#define FUSED_WAVES_PER_EU __attribute__((amdgpu_waves_per_eu(2, 8)))
constexpr int HOP_FUSED = 2;

// sum_m W[m] * x[m] over four terms (lazily for BabyBear^4 weights on
// BabyBear values: four products per limb stay below 2 p 2^32).
template <class FC, class FT>
__device__ __forceinline__ typename FC::T mix4(const typename FC::T (&W)[4],
                                               const typename FT::T (&x)[4]) {
  typename FC::T s = MixMul<FC, FT>::go(W[0], x[0]);
#pragma unroll
  for (int m = 1; m < 4; ++m) s = FC::add(s, MixMul<FC, FT>::go(W[m], x[m]));
  return s;
}
template <>
__device__ __forceinline__ bb4 mix4<BB4, BB>(const bb4 (&W)[4], const uint32_t (&x)[4]) {
  uint32_t o[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    uint64_t s = (uint64_t)W[0].c[c] * x[0];
#pragma unroll
    for (int m = 1; m < 4; ++m) s += (uint64_t)W[m].c[c] * x[m];
    o[c] = bb_redc(bb_pre(s));
  }
  return bb4_make(o[0], o[1], o[2], o[3]);
}

// Round 2's entry from its four input entries at offsets 0, N/2, N/4, 3N/4
// (x), folded at r0 then r1: two folds of the challenge field, or for
// BabyBear entries under BabyBear^4 one lazy four-term sum with the weights
// W = ((1-r1)(1-r0), (1-r1) r0, r1 (1-r0), r1 r0).
template <class FC, class FT>
struct Fold2 {
  static __device__ __forceinline__ typename FC::T go(const typename FC::T (&W)[4],
                                                      typename FC::T r0, typename FC::T r1,
                                                      const typename FT::T (&x)[4]) {
    using T = typename FC::T;
    const T x0 = Lift<FT, FC>::go(x[0]), x1 = Lift<FT, FC>::go(x[1]);
    const T x2 = Lift<FT, FC>::go(x[2]), x3 = Lift<FT, FC>::go(x[3]);
    const T a = FC::add(x0, FC::mul(r0, FC::sub(x1, x0)));
    const T b = FC::add(x2, FC::mul(r0, FC::sub(x3, x2)));
    return FC::add(a, FC::mul(r1, FC::sub(b, a)));
  }
};
template <>
struct Fold2<BB4, BB> {
  static __device__ __forceinline__ bb4 go(const bb4 (&W)[4], bb4, bb4, const uint32_t (&x)[4]) {
    return mix4<BB4, BB>(W, x);
  }
};

// fused01_kernel's sums: eq-weighted (lazily for BabyBear^4 weights on
// BabyBear values), or without eq plain base-field sums lifted at the end.
template <class FC, class FT, bool EQ>
struct FusedSum : WeightedSum<FC, FT> {};
template <class FC, class FT>
struct FusedSum<FC, FT, false> {
  typename FT::T s;
  __device__ __forceinline__ void init() { s = FT::zero(); }
  __device__ __forceinline__ void add(typename FC::T, typename FT::T x, bool = true) {
    s = FT::add(s, x);
  }
  __device__ __forceinline__ typename FC::T get() const { return Lift<FT, FC>::go(s); }
};

// Eq weight of pair j of a round: with `fac`, the per-thread factor wr is
// applied to the sums at the end and the pair takes wl alone.
template <class FC>
__device__ __forceinline__ typename FC::T eq_w(const typename FC::T* wl, const typename FC::T* wr,
                                               uint32_t nx, bool fac, size_t j) {
  if (fac) return wl[j >> nx];
  if (wl) return FC::mul(wl[j >> nx], wr[j & (((size_t)1 << nx) - 1)]);
  return wr[j];
}

// fused01_kernel's streaming loop: groups g = start + i stride < N/4, U a
// step, the next step's entries (and, with FAC, its eq weights: wl alone,
// the per-thread wr factor applied by the caller) loaded before this step's
// arithmetic. Group g: input entries g + m N/4 (m = 0..3) of every table. Its
// four entries of table q span the (x_0, x_1) square of round 1's pair g, so
// round 1's factor at point t folded at r_0 = s is
//   T_q(s, t) = A_q + s B_q,  A_q = x0 + t (x1 - x0),  B_q = (x2 - x0) + t ((x3 - x1) - (x2 - x0)),
// and csum[t][i] sums the coefficient of s^i of prod_q T_q(s, t) weighted by
// round 1's eq weight of pair g. Round 0's node sums follow from these: its
// pair (x_1, g) has eq weight eq(w_1, x_1) times round 1's, so
//   round0(s) = sum_i s^i (eq(w_1, 0) csum[0][i] + eq(w_1, 1) csum[1][i])
// (both weights 1 without eq), which the finisher forms. Groups a step: as
// many as keep the entries of two steps and the (K+1)^2 sums in registers.
template <int K>
struct FusedU {
  static constexpr int value = K <= 2 ? 4 : K == 3 ? 2 : 1;
};
template <class FT, class FC, int K, bool EQ, bool FAC>
__device__ __forceinline__ void fused01_stream(const StageArgs<FT, FC>& a, uint32_t start,
                                               uint32_t stride, const typename FC::T* wl1,
                                               const typename FC::T* wr1, uint32_t x1,
                                               FusedSum<FC, FT, EQ> (&csum)[K + 1][K + 1]) {
  using T = typename FC::T;
  using S = typename FT::T;
  constexpr int NA = K + 1;
  constexpr int U = FusedU<K>::value;
  constexpr bool PW = EQ && FAC;  // weights prefetched with the entries
  constexpr int UW = PW ? U : 1;
  // Group indices fit 32 bits, and so do their byte offsets into a quarter
  // table (n <= 31): each stream is a uniform base and a 32-bit byte offset.
  const uint32_t q4 = 1u << (a.n - 2);
  const char* base[K][4];
#pragma unroll
  for (int q = 0; q < K; ++q)
#pragma unroll
    for (int m = 0; m < 4; ++m)
      base[q][m] = reinterpret_cast<const char*>(a.in + (size_t)q * a.in_stride + (size_t)m * q4);
  const uint32_t step = (uint32_t)U * stride;
  // Loads of the U groups from g0 (indices past the round clamped into it,
  // so that every step issues the same loads).
  auto load = [&](uint32_t g0, S (&x)[U][K][4], T (&wc)[UW]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t g = min(g0 + (uint32_t)u * stride, q4 - 1u);
      const uint32_t off = g * (uint32_t)sizeof(S);
#pragma unroll
      for (int q = 0; q < K; ++q)
#pragma unroll
        for (int m = 0; m < 4; ++m) x[u][q][m] = *reinterpret_cast<const S*>(base[q][m] + off);
      if constexpr (PW) wc[u] = wl1[g >> x1];
    }
  };
  // Group g from its entries e with weight w; `fold`: whether the lazy sums
  // fold after it (every second group).
  auto group = [&](const S (&e)[K][4], T w, bool fold) {
    S A[K], B[K], dA[K], dB[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      A[q] = e[q][0];
      B[q] = FT::sub(e[q][2], e[q][0]);
      dA[q] = FT::sub(e[q][1], e[q][0]);
      dB[q] = FT::sub(FT::sub(e[q][3], e[q][1]), B[q]);
    }
#pragma unroll
    for (int t = 0; t < NA; ++t) {
      S p[K + 1];
      la_poly<FT, K>(A, B, p);
#pragma unroll
      for (int i = 0; i < NA; ++i) csum[t][i].add(w, p[i], fold);
      if (t < K) {
#pragma unroll
        for (int q = 0; q < K; ++q) {
          A[q] = FT::add(A[q], dA[q]);
          B[q] = FT::add(B[q], dB[q]);
        }
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  };
  auto weight = [&](uint32_t g) -> T {
    if constexpr (EQ) return eq_w<FC>(wl1, wr1, x1, false, g);
    return FC::one();
  };
  // Steps whose U groups are all in the round (start < stride).
  const uint32_t nfull = q4 / step;
  uint32_t g0 = start;
  S cur[U][K][4];
  T cc[UW];
  if (nfull > 0) load(g0, cur, cc);
  for (uint32_t k = 0; k < nfull; ++k, g0 += step) {
    S nxt[U][K][4];
    T nc[UW];
    load(g0 + step, nxt, nc);
#pragma unroll
    for (int u = 0; u < U; ++u)
      group(cur[u], PW ? cc[u] : weight(g0 + (uint32_t)u * stride), (u & 1) || u == U - 1);
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
      for (int q = 0; q < K; ++q)
#pragma unroll
        for (int m = 0; m < 4; ++m) cur[u][q][m] = nxt[u][q][m];
      if constexpr (PW) cc[u] = nc[u];
    }
  }
  // The last, partial step.
  if (g0 < q4) {
    load(g0, cur, cc);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t g = g0 + (uint32_t)u * stride;
      if (g < q4) group(cur[u], PW ? cc[u] : weight(g), true);
    }
  }
}

// The entries of one 16-byte vector (16 / sizeof(S) consecutive elements).
template <class S>
__device__ __forceinline__ void unpack16(const uint4& v, S (&o)[16 / sizeof(S)]) {
  static_assert(16 % sizeof(S) == 0, "element size");
  __builtin_memcpy(&o[0], &v, 16);
}

// fused01_kernel's streaming loop over 16-byte vectors: a lane loads VG =
// 16 / sizeof(S) consecutive groups of every stream at once (a wave reads
// 1 KiB a stream per load). Round 1's eq weight of group g is L[g >> x] R[g &
// (2^x - 1)] (or R[g] alone, or none); the groups are split into H = 2^(n-2-x)
// blocks of equal L, and P = T / H threads share a block, each taking chunks
// cl0, cl0 + P, ... of it, so that L is a per-thread factor (applied to the
// sums at the end, as `lfac`) and R the per-group weight, read with the
// entries (R is 2^x elements: cache-resident).
template <class FT, class FC, int K, bool EQ>
__device__ __forceinline__ void fused01_stream_vec(const StageArgs<FT, FC>& a, uint32_t t,
                                                   uint32_t P, uint32_t xr, const typename FC::T* R,
                                                   FusedSum<FC, FT, EQ> (&csum)[K + 1][K + 1]) {
  using T = typename FC::T;
  using S = typename FT::T;
  constexpr int NA = K + 1;
  constexpr int VG = 16 / (int)sizeof(S);
  constexpr uint32_t LVG = VG == 4 ? 2u : VG == 2 ? 1u : 0u;
  constexpr int NW = EQ ? VG : 1;
  const uint32_t q4 = 1u << (a.n - 2);
  const uint32_t hb = 1u << (xr - LVG);  // chunks of one block of equal L
  const uint32_t h = t / P, cl0 = t % P, ns = hb / P;
  const char* base[K][4];
#pragma unroll
  for (int q = 0; q < K; ++q)
#pragma unroll
    for (int m = 0; m < 4; ++m)
      base[q][m] = reinterpret_cast<const char*>(a.in + (size_t)q * a.in_stride + (size_t)m * q4);
  // Step s's chunk (steps past the last reload the last: in bounds).
  auto load = [&](uint32_t s, uint4 (&x)[K][4], T (&w)[NW]) {
    const uint32_t cl = cl0 + min(s, ns - 1u) * P;
    const uint32_t off = ((h * hb + cl) << 4);
#pragma unroll
    for (int q = 0; q < K; ++q)
#pragma unroll
      for (int m = 0; m < 4; ++m) x[q][m] = *reinterpret_cast<const uint4*>(base[q][m] + off);
    if constexpr (EQ) {
      const T* r = R + (cl << LVG);
#pragma unroll
      for (int v = 0; v < VG; ++v) w[v] = r[v];
    }
  };
  auto group = [&](const S (&e)[K][4], T w, bool fold) {
    S A[K], B[K], dA[K], dB[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      A[q] = e[q][0];
      B[q] = FT::sub(e[q][2], e[q][0]);
      dA[q] = FT::sub(e[q][1], e[q][0]);
      dB[q] = FT::sub(FT::sub(e[q][3], e[q][1]), B[q]);
    }
#pragma unroll
    for (int tt = 0; tt < NA; ++tt) {
      S p[K + 1];
      la_poly<FT, K>(A, B, p);
#pragma unroll
      for (int i = 0; i < NA; ++i) csum[tt][i].add(w, p[i], fold);
      if (tt < K) {
#pragma unroll
        for (int q = 0; q < K; ++q) {
          A[q] = FT::add(A[q], dA[q]);
          B[q] = FT::add(B[q], dB[q]);
        }
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  };
  uint4 cur[K][4];
  T cw[NW];
  load(0, cur, cw);
  for (uint32_t s = 0; s < ns; ++s) {
    uint4 nxt[K][4];
    T nw[NW];
    load(s + 1, nxt, nw);
    S el[K][4][VG];
#pragma unroll
    for (int q = 0; q < K; ++q)
#pragma unroll
      for (int m = 0; m < 4; ++m) unpack16<S>(cur[q][m], el[q][m]);
#pragma unroll
    for (int v = 0; v < VG; ++v) {
      S e[K][4];
#pragma unroll
      for (int q = 0; q < K; ++q)
#pragma unroll
        for (int m = 0; m < 4; ++m) e[q][m] = el[q][m][v];
      group(e, EQ ? cw[EQ ? v : 0] : FC::one(), (v & 1) || v == VG - 1);
    }
#pragma unroll
    for (int q = 0; q < K; ++q)
#pragma unroll
      for (int m = 0; m < 4; ++m) cur[q][m] = nxt[q][m];
#pragma unroll
    for (int v = 0; v < NW; ++v) cw[v] = nw[v];
  }
}

template <class FC, int K>
struct Fused01Lds {
  static constexpr int NA = K + 1, NC = NA * NA;
  RoundLds<FC, NC> round;
  HopLds warm;
};

// The finisher of rounds 0 and 1 (one block, after every streaming block of
// the proof has added its coefficient sums): round 0's message from round
// 1's coefficient sums, its hop, round 1's node sums at r_0, its hop. The
// proof state must already be in sh.sm (fin_load, fin_eq of round 0).
template <class FT, class FC, int K, bool EQ, int HOP>
__device__ __forceinline__ void fused01_finish(const StageArgs<FT, FC>& a,
                                               RoundLds<FC, (K + 1) * (K + 1)>& sh) {
  using T = typename FC::T;
  constexpr int NA = K + 1, NC = NA * NA;
  constexpr int W = SlotOps<FC>::W;
  const uint32_t tid = threadIdx.x;
  Ctl<FC>* ctl = a.ctl;
  StageArgs<FT, FC> af = a;
  af.final_out = nullptr;
  static_assert(sizeof(sh.sum64) == (PASS_THREADS / LA_SLOTW) * LA_SLOTW * sizeof(uint64_t),
                "slot scratch shapes");
  la_slot_sum<FC, NC>(la_rows(a.slots, 1), NA * W,
                      reinterpret_cast<uint64_t (*)[LA_SLOTW]>(&sh.sum64[0][0]), sh.out);
  // Round 0's node sums from round 1's coefficients (fused01_stream):
  // node t = sum_i t^i (e0 c[0][i] + e1 c[1][i]), e_b = eq(w_1, b).
  if (tid < (uint32_t)NA) {
    T e0 = FC::one(), e1 = FC::one();
    if constexpr (EQ) {
      e0 = a.eq_ab[3];                   // 1 - w_1
      e1 = FC::add(a.eq_ab[2], a.eq_ab[3]);  // w_1
    }
    T v = FC::zero();
#pragma unroll
    for (int i = K; i >= 0; --i) {
      const T u = FC::add(FC::mul(e0, sh.out[i]), FC::mul(e1, sh.out[NA + i]));
      T vt = FC::zero();
      for (uint32_t j = 0; j < tid; ++j) vt = FC::add(vt, v);  // v * t, t <= K
      v = FC::add(vt, u);
    }
    sh.lds[tid] = v;
  }
  __syncthreads();
  if (tid == 0) trace_mark(a.trace, 0, 2);
  if (tid < 64)
    finish_round<FT, FC, K, EQ, HOP>(af, 0, false, sh.lds, sh.sm, a.work, a.in,
                                           a.in_stride, true);
  // Round 1's node sums: its coefficients at r_0.
  if (tid < (uint32_t)NA) {
    const T r = sh.sm.r;
    T v = sh.out[tid * NA + K];
#pragma unroll
    for (int i = K - 1; i >= 0; --i) v = FC::add(sh.out[tid * NA + i], FC::mul(r, v));
    sh.lds[tid] = v;
  }
  __syncthreads();
  if (tid < (uint32_t)NA) sh.out[tid] = sh.lds[tid];
  if (tid < 64) fin_eq<FT, FC, EQ>(sh.sm, a, 1);
  __syncthreads();
  if (tid == 0) trace_mark(a.trace, 1, 2);
  if (tid < 64) {
    finish_round<FT, FC, K, EQ, HOP>(af, 1, false, sh.out, sh.sm, a.work, a.in,
                                           a.in_stride, true);
    fin_store(sh.sm, ctl);
  }
}

// One streaming block of rounds 0 and 1: block `blk` of `nblk` sums its
// groups' (K+1)^2 coefficients and adds them to the look-ahead rows of round
// 1 (the caller then arrives).
template <class FT, class FC, int K, bool EQ>
__device__ __forceinline__ void fused01_stream_block(const StageArgs<FT, FC>& a, uint32_t blk,
                                                     uint32_t nblk,
                                                     RoundLds<FC, (K + 1) * (K + 1)>& sh) {
  using T = typename FC::T;
  using S = typename FT::T;
  constexpr int NA = K + 1, NC = NA * NA;
  constexpr int W = SlotOps<FC>::W;
  const uint32_t tid = threadIdx.x, n = a.n;
  Ctl<FC>* ctl = a.ctl;
  const uint32_t q4 = 1u << (n - 2);  // round 1's pairs
  const uint32_t start = blk * PASS_THREADS + tid, stride = nblk * PASS_THREADS;
  const T* wl1 = nullptr;
  const T* wr1 = nullptr;
  uint32_t x1 = 0;
  bool fac = false;
  if constexpr (EQ) {
    select_eq<FC>(1, n, a.eq_l_fam, a.eq_r_fam, &ctl->one, wl1, wr1, x1);
    const uint32_t mask = (1u << x1) - 1u;
    fac = wl1 && (stride & mask) == 0;
  }
  FusedSum<FC, FT, EQ> csum[NA][NA];
#pragma unroll
  for (int t = 0; t < NA; ++t)
#pragma unroll
    for (int i = 0; i < NA; ++i) csum[t][i].init();
  // The 16-byte-vector loop when the groups split into blocks of equal L
  // (one per P threads, P a power of two no larger than a block's chunks).
  constexpr uint32_t LVG = sizeof(S) == 4 ? 2u : sizeof(S) == 8 ? 1u : 0u;
  const uint32_t xr = EQ && wl1 ? x1 : n - 2;
  const uint32_t nch = q4 >> LVG;
  const uint32_t H = xr >= LVG ? nch >> (xr - LVG) : 0u;
  const bool vec = H != 0 && H <= stride && (stride / H) <= (1u << (xr - LVG));
  T fac_w = FC::one();
  if (vec) {
    const uint32_t P = stride / H;
    fused01_stream_vec<FT, FC, K, EQ>(a, start, P, xr, wr1, csum);
    if (EQ && wl1) fac_w = wl1[start / P];
  } else if (fac) {
    fused01_stream<FT, FC, K, EQ, true>(a, start, stride, wl1, wr1, x1, csum);
    fac_w = start < q4 ? wr1[start & ((1u << x1) - 1u)] : FC::one();
  } else {
    fused01_stream<FT, FC, K, EQ, false>(a, start, stride, wl1, wr1, x1, csum);
  }
  T c[NC];
#pragma unroll
  for (int t = 0; t < NA; ++t) {
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      c[t * NA + i] = csum[t][i].get();
      if (EQ) c[t * NA + i] = FC::mul(c[t * NA + i], fac_w);
    }
  }
  block_reduce<FC, NC, PASS_THREADS>(c, sh.lds, sh.out);
  la_slot_add<FC, NC>(la_rows(a.slots, 1), blk, NA * W, sh.out);
}

// Rounds 0 and 1 of one proof as 1 + nblk blocks (round 1 < n - 1).
template <class FT, class FC, int K, bool EQ>
__global__ __launch_bounds__(PASS_THREADS) FUSED_WAVES_PER_EU void fused01_kernel(
    StageArgs<FT, FC> a) {
  using T = typename FC::T;
  using S = typename FT::T;
  constexpr int NA = K + 1, NC = NA * NA;
  constexpr int NPTS = K + 1 + (EQ ? 1 : 0);
  constexpr int W = SlotOps<FC>::W;
  __shared__ Fused01Lds<FC, K> ps;
  RoundLds<FC, NC>& sh = ps.round;
  const uint32_t tid = threadIdx.x, nblk = gridDim.x - 1, n = a.n;
  Ctl<FC>* ctl = a.ctl;
  if (blockIdx.x == 0) {
    if (tid < 64) {
      fin_eq<FT, FC, EQ>(sh.sm, a, 0);
      fin_load(sh.sm, ctl);
      const uint32_t* from = reinterpret_cast<const uint32_t*>(static_cast<HopLds*>(&sh.sm));
      uint32_t* to = reinterpret_cast<uint32_t*>(&ps.warm);
      for (uint32_t i = tid; i < sizeof(HopLds) / 4; i += 64u) to[i] = from[i];
      wave_sync();
      fs_hop<HOP_FUSED>((ZS_LDS HopLds*)&ps.warm, (const ZS_LDS uint8_t*)sh.sm.msg,
                        (uint32_t)(NPTS * sizeof(T)), ctl->item_bytes, ctl->limbs, ctl->tr_kind);
    }
    if (tid == 0) sh.is_last = chain_wait(&ctl->counter, nblk, &ctl->fault) ? 1u : 0u;
    __syncthreads();
    if (!sh.is_last) return;
    fused01_finish<FT, FC, K, EQ, HOP_FUSED>(a, sh);
    if (tid == 0) __hip_atomic_store(&ctl->counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  fused01_stream_block<FT, FC, K, EQ>(a, blockIdx.x - 1, nblk, sh);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (tid == 0) __hip_atomic_fetch_add(&ctl->counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One streaming block of round 2 from the input: block `blk` of `nblk`
// folds its pairs' input entries at r_0 and r_1, writes them to `work`, and
// adds its node sums to the slot rows (the caller then arrives).
template <class FT, class FC, int K, bool EQ>
__device__ __forceinline__ void fold2_stream_block(const StageArgs<FT, FC>& a, uint32_t blk,
                                                   uint32_t nblk, RoundLds<FC, K + 1>& sh) {
  using T = typename FC::T;
  using S = typename FT::T;
  constexpr int NA = K + 1;
  const uint32_t tid = threadIdx.x, n = a.n;
  Ctl<FC>* ctl = a.ctl;
  const size_t q8 = (size_t)1 << (n - 3);  // round 2's pairs
  const size_t start = (size_t)blk * PASS_THREADS + tid, stride = (size_t)nblk * PASS_THREADS;
  // Weights of the two folds.
  T Wt[4];
  const T r0 = el_from_words<FC>(reinterpret_cast<const uint32_t*>(a.point));
  const T r1 = ctl->r;
  {
    const T u0 = FC::sub(FC::one(), r0), u1 = FC::sub(FC::one(), r1);
    Wt[0] = FC::mul(u1, u0);  // offset 0
    Wt[1] = FC::mul(u1, r0);  // offset N/2
    Wt[2] = FC::mul(r1, u0);  // offset N/4
    Wt[3] = FC::mul(r1, r0);  // offset 3N/4
  }
  const T* wl = nullptr;
  const T* wr = nullptr;
  uint32_t nx = 0;
  bool fac = false;
  if constexpr (EQ) {
    select_eq<FC>(2, n, a.eq_l_fam, a.eq_r_fam, &ctl->one, wl, wr, nx);
    fac = wl && (stride & (((size_t)1 << nx) - 1)) == 0;
  }
  T acc[NA];
#pragma unroll
  for (int t = 0; t < NA; ++t) acc[t] = FC::zero();
  // Pair j: input entries j + m N/8, m = 0..7, of every table, and with
  // `fac` its eq weight (wl alone; wr is this thread's factor).
  auto load = [&](size_t j, S (&x)[K][8], T& wv) {
    j = j < q8 ? j : q8 - 1;
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const S* s = a.in + (size_t)q * a.in_stride;
#pragma unroll
      for (int m = 0; m < 8; ++m) x[q][m] = s[j + (size_t)m * q8];
    }
    if (EQ && fac) wv = wl[j >> nx];
  };
  S cur[K][8];
  T cw = FC::one();
  if (start < q8) load(start, cur, cw);
  for (size_t j = start; j < q8; j += stride) {
    T w = cw;
    if (EQ && !fac) w = eq_w<FC>(wl, wr, nx, false, j);
    S nxt[K][8];
    T nw = FC::one();
    load(j + stride, nxt, nw);
    T lo[K], df[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      T* d = a.work + (size_t)q * a.work_stride;
      T v[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        // Offsets 0, N/2, N/4, 3N/4 of entry j + h N/8: m = h, h + 4, h + 2, h + 6.
        const S x4[4] = {cur[q][h], cur[q][h + 4], cur[q][h + 2], cur[q][h + 6]};
        v[h] = Fold2<FC, FT>::go(Wt, r0, r1, x4);
        d[j + (size_t)h * q8] = v[h];
      }
      lo[q] = v[0];
      df[q] = FC::sub(v[1], v[0]);
    }
    if constexpr (EQ) {
      const T hi0 = FC::add(lo[0], df[0]);
      lo[0] = FC::mul(w, lo[0]);
      df[0] = FC::sub(FC::mul(w, hi0), lo[0]);
    }
#pragma unroll
    for (int t = 0; t < NA; ++t) {
      T prod = lo[0];
#pragma unroll
      for (int q = 1; q < K; ++q) prod = FC::mul(prod, lo[q]);
      acc[t] = FC::add(acc[t], prod);
      if (t < K) {
#pragma unroll
        for (int q = 0; q < K; ++q) lo[q] = FC::add(lo[q], df[q]);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < K; ++q)
#pragma unroll
      for (int m = 0; m < 8; ++m) cur[q][m] = nxt[q][m];
    cw = nw;
  }
  if (EQ && fac && start < q8) {
    const T wrc = wr[start & (((size_t)1 << nx) - 1)];
#pragma unroll
    for (int t = 0; t < NA; ++t) acc[t] = FC::mul(acc[t], wrc);
  }
  block_reduce<FC, NA, PASS_THREADS>(acc, sh.lds, sh.out);
  slot_row_add<FC, NA>(a.slots, blk, sh.out);
}

// Round 2 of one proof from the input as 1 + nblk blocks: pair j's table
// entries T2[j + h N/8] (h = 0, 1) are the input entries at j + h N/8 +
// {0, N/2, N/4, 3N/4} folded at r_0 and r_1 (Fold2); they are written to
// `work` and evaluated; block 0 finishes round 2 as in pass_kernel. r_0 is
// read from the raw point (the control block holds r_1).
template <class FT, class FC, int K, bool EQ>
__global__ __launch_bounds__(PASS_THREADS) PASS_WAVES_PER_EU void fold2_kernel(
    StageArgs<FT, FC> a, uint32_t need_claim) {
  using T = typename FC::T;
  using S = typename FT::T;
  constexpr int NA = K + 1;
  constexpr int NPTS = K + 1 + (EQ ? 1 : 0);
  __shared__ PassLds<FC, NA> ps;
  RoundLds<FC, NA>& sh = ps.round;
  const uint32_t tid = threadIdx.x, nblk = gridDim.x - 1, n = a.n;
  Ctl<FC>* ctl = a.ctl;
  if (blockIdx.x == 0) {
    if (tid < 64) {
      fin_eq<FT, FC, EQ>(sh.sm, a, 2);
      fin_load(sh.sm, ctl);
      const uint32_t* from = reinterpret_cast<const uint32_t*>(static_cast<HopLds*>(&sh.sm));
      uint32_t* to = reinterpret_cast<uint32_t*>(&ps.warm);
      for (uint32_t i = tid; i < sizeof(HopLds) / 4; i += 64u) to[i] = from[i];
      wave_sync();
      fs_hop<HOP_PASS>((ZS_LDS HopLds*)&ps.warm, (const ZS_LDS uint8_t*)sh.sm.msg,
                       (uint32_t)(NPTS * sizeof(T)), ctl->item_bytes, ctl->limbs, ctl->tr_kind);
    }
    if (tid == 0) sh.is_last = chain_wait(&ctl->counter, nblk, &ctl->fault) ? 1u : 0u;
    __syncthreads();
    if (!sh.is_last) return;
    if (tid == 0) trace_mark(a.trace, 2, 2);
    slot_rows_sum<FC, NA>(a.slots, sh);
    if (tid < 64) {
      finish_round<FT, FC, K, EQ, HOP_PASS>(a, 2, need_claim != 0, sh.out, sh.sm, a.work, a.in,
                                            a.work_stride, false);
      fin_store(sh.sm, ctl);
    }
    if (tid == 0) __hip_atomic_store(&ctl->counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  fold2_stream_block<FT, FC, K, EQ>(a, blockIdx.x - 1, nblk, sh);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (tid == 0) __hip_atomic_fetch_add(&ctl->counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------ batched steps
// One step over many proofs: each item is one round of one proof. Item i owns
// blocks [block_off, block_off + blocks) of the launch; a tail item is one
// block that runs a run of its proof's small rounds. Items are sorted by
// block_off.
struct BatchItem {
  uint32_t inst;        // index into the step's StageArgs array
  uint32_t blocks;      // power of two
  uint32_t block_off;
  uint32_t round_tail;  // a pass: round; a tail: tail_item(round, end), rounds [round, end)
};
constexpr uint32_t TAIL_FLAG = 0x80000000u;
constexpr uint32_t ROUND_MASK = 0xffffu;
constexpr uint32_t TAIL_END_SHIFT = 16;
ZS_HD constexpr uint32_t tail_item(uint32_t round, uint32_t end) {
  return round | end << TAIL_END_SHIFT | TAIL_FLAG;
}

// Item of this block: the number of items whose block_off <= blockIdx.x,
// minus one, counted by the whole block in chunks of PASS_THREADS items.
__device__ __forceinline__ uint32_t find_item(const BatchItem* items, uint32_t nitems) {
  uint32_t below = 0;
  for (uint32_t base = 0; base < nitems; base += PASS_THREADS) {
    const uint32_t i = base + threadIdx.x;
    below += __syncthreads_count(i < nitems && items[i].block_off <= blockIdx.x);
  }
  return below - 1;
}

template <class FT, class FC, int NA>
union BatchLds {
  RoundLds<FC, NA> round;
  TailLds<FC, NA, PASS_THREADS> tail;
};

// A tail item of a batch: rounds [round, end) of one proof in this block.
// Out of line, so that the tail's registers do not weigh on the streaming
// loop of the pass items.
template <class FT, class FC, int K, bool EQ>
static __device__ __attribute__((noinline)) void batch_tail(
    const StageArgs<FT, FC>* __restrict__ args, uint32_t round, uint32_t end,
    ZS_LDS TailLds<FC, K + 1, PASS_THREADS>* sh) {
  const StageArgs<FT, FC> a = *args;
  tail_rounds<FT, FC, K, EQ, PASS_THREADS>(a, round, end,
                                           *(TailLds<FC, K + 1, PASS_THREADS>*)sh);
}

// RT: the round type of every pass item of the launch: 0 = round 0 (input
// tables, no fold), 1 = round 1 (fold of the input tables), 2 = later rounds
// (fold in the challenge field). Tail items may ride with any of them.
template <class FT, class FC, int K, bool EQ, int RT>
__global__ __launch_bounds__(PASS_THREADS) PASS_WAVES_PER_EU void batch_kernel(
    const StageArgs<FT, FC>* __restrict__ args, const BatchItem* __restrict__ items,
    uint32_t nitems) {
  __shared__ BatchLds<FT, FC, K + 1> sh;
  const uint32_t idx = find_item(items, nitems);
  const BatchItem it = items[idx];
  const StageArgs<FT, FC> a = args[it.inst];
  const uint32_t round = it.round_tail & ROUND_MASK;
  if (it.round_tail & TAIL_FLAG) {
    batch_tail<FT, FC, K, EQ>(args + it.inst, round,
                              (it.round_tail & ~TAIL_FLAG) >> TAIL_END_SHIFT,
                              (ZS_LDS TailLds<FC, K + 1, PASS_THREADS>*)&sh.tail);
    return;
  }
  const uint32_t blk = blockIdx.x - it.block_off;
  const uint32_t L = 1u << (a.n - round);
  const bool need_claim = round + 1 == a.n;
  static_assert(RT >= 0 && RT <= 2, "round type");
  if constexpr (RT == 0) {
    round_block<FT, FT, FC, K, EQ, false>(a, a.in, a.in_stride, L, round, need_claim, blk,
                                          it.blocks, sh.round);
  } else if constexpr (RT == 1) {
    round_block<FT, FT, FC, K, EQ, true>(a, a.in, a.in_stride, L, round, need_claim, blk,
                                         it.blocks, sh.round);
  } else {
    round_block<FT, FC, FC, K, EQ, true>(a, a.work, a.work_stride, L, round, need_claim, blk,
                                         it.blocks, sh.round);
  }
}

// Rounds 0 and 1 of one batch member from the input as `nblk` blocks of a
// step (fused01_kernel's work): every block streams its groups and adds its
// coefficient sums; the last block to take a ticket finishes both rounds.
template <class FT, class FC, int K, bool EQ>
__device__ __forceinline__ void fused01_block(const StageArgs<FT, FC>& a, uint32_t blk,
                                              uint32_t nblk,
                                              RoundLds<FC, (K + 1) * (K + 1)>& sh) {
  fused01_stream_block<FT, FC, K, EQ>(a, blk, nblk, sh);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t prev =
        __hip_atomic_fetch_add(&a.ctl->counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sh.is_last = (prev == nblk - 1) ? 1u : 0u;
  }
  __syncthreads();
  if (!sh.is_last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  if (threadIdx.x < 64) {
    fin_eq<FT, FC, EQ>(sh.sm, a, 0);
    fin_load(sh.sm, a.ctl);
  }
  __syncthreads();
  fused01_finish<FT, FC, K, EQ, HOP_PASS>(a, sh);
  if (threadIdx.x == 0)
    __hip_atomic_store(&a.ctl->counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Round 2 of one batch member from the input as `nblk` blocks of a step
// (fold2_kernel's work); the last block to take a ticket finishes round 2.
template <class FT, class FC, int K, bool EQ>
__device__ __forceinline__ void fold2_block(const StageArgs<FT, FC>& a, bool need_claim,
                                            uint32_t blk, uint32_t nblk,
                                            RoundLds<FC, K + 1>& sh) {
  fold2_stream_block<FT, FC, K, EQ>(a, blk, nblk, sh);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t prev =
        __hip_atomic_fetch_add(&a.ctl->counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sh.is_last = (prev == nblk - 1) ? 1u : 0u;
  }
  __syncthreads();
  if (!sh.is_last) return;
  if (threadIdx.x < 64) {
    fin_eq<FT, FC, EQ>(sh.sm, a, 2);
    fin_load(sh.sm, a.ctl);
  }
  slot_rows_sum<FC, K + 1>(a.slots, sh);
  if (threadIdx.x < 64) {
    finish_round<FT, FC, K, EQ, HOP_PASS>(a, 2, need_claim, sh.out, sh.sm, a.work, a.in,
                                          a.work_stride, false);
    fin_store(sh.sm, a.ctl);
  }
  if (threadIdx.x == 0)
    __hip_atomic_store(&a.ctl->counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Batch steps of members whose first rounds come from the input: rounds 0-1
// (FIRST = 0) or round 2 (FIRST = 2) of every item (no tail items).
template <class FT, class FC, int K, bool EQ>
__global__ __launch_bounds__(PASS_THREADS) FUSED_WAVES_PER_EU void batch01_kernel(
    const StageArgs<FT, FC>* __restrict__ args, const BatchItem* __restrict__ items,
    uint32_t nitems) {
  __shared__ RoundLds<FC, (K + 1) * (K + 1)> sh;
  const uint32_t idx = find_item(items, nitems);
  const BatchItem it = items[idx];
  const StageArgs<FT, FC> a = args[it.inst];
  fused01_block<FT, FC, K, EQ>(a, blockIdx.x - it.block_off, it.blocks, sh);
}

template <class FT, class FC, int K, bool EQ>
__global__ __launch_bounds__(PASS_THREADS) PASS_WAVES_PER_EU void batch2_kernel(
    const StageArgs<FT, FC>* __restrict__ args, const BatchItem* __restrict__ items,
    uint32_t nitems) {
  __shared__ RoundLds<FC, K + 1> sh;
  const uint32_t idx = find_item(items, nitems);
  const BatchItem it = items[idx];
  const StageArgs<FT, FC> a = args[it.inst];
  fold2_block<FT, FC, K, EQ>(a, 3 == a.n, blockIdx.x - it.block_off, it.blocks, sh);
}

// Suffix eq families of the two halves of the eq point (compute_eq_evaluations,
// eq_poly.py:46-51): member m of the family of v[0:h] is eq(v[h-1-m:], .) over
// {0,1}^(m+1) with v[h-1-m] on the MSB, stored at fam_offset(m). The reference
// grows each member from the previous one ([prev (1 - c), prev c]); here every
// entry is its product over coordinates, prod_i (bit_i ? c_i : 1 - c_i),
// computed independently by one thread (the same field value, exactly).
template <class FC>
__device__ __forceinline__ void eq_family_entry(const typename FC::T* v, uint32_t h_l,
                                                typename FC::T* fam_l, uint32_t h_r,
                                                typename FC::T* fam_r, size_t g) {
  using T = typename FC::T;
  const size_t n_l = fam_offset(h_l), n_r = fam_offset(h_r);
  if (g >= n_l + n_r) return;
  const bool left = g < n_l;
  const size_t idx = left ? g : g - n_l;
  const uint32_t h = left ? h_l : h_r;
  const T* vv = left ? v : v + h_l;
  T* fam = left ? fam_l : fam_r;
  const uint32_t m = 31u - __builtin_clz((uint32_t)(idx + 2)) - 1u;  // fam_offset(m) <= idx
  const size_t x = idx - fam_offset(m);
  const T one = FC::one();
  T acc = one;
  for (uint32_t i = 0; i <= m; ++i) {
    const T c = vv[h - 1 - m + i];
    const bool bit = (x >> (m - i)) & 1u;
    acc = FC::mul(acc, bit ? c : FC::sub(one, c));
  }
  fam[idx] = acc;
}

// The point's coordinates staged in LDS first: each entry reads up to n of
// them in turn, and from global memory every read waited a cache latency.
constexpr uint32_t EQ_STAGE_MAX = 64;
template <class FC>
__device__ __forceinline__ const typename FC::T* eq_stage(const typename FC::T* v, uint32_t n,
                                                          typename FC::T* vs) {
  if (n > EQ_STAGE_MAX) return v;
  if (threadIdx.x < n) vs[threadIdx.x] = v[threadIdx.x];
  __syncthreads();
  return vs;
}

template <class FC>
__global__ __launch_bounds__(256) void eq_family_kernel(const typename FC::T* v, uint32_t h_l,
                                                        typename FC::T* fam_l, uint32_t h_r,
                                                        typename FC::T* fam_r) {
  __shared__ typename FC::T vs[EQ_STAGE_MAX];
  const typename FC::T* w = eq_stage<FC>(v, h_l + h_r, vs);
  eq_family_entry<FC>(w, h_l, fam_l, h_r, fam_r, (size_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// The eq families of many proofs in one launch: job i owns blocks
// [block_off, next block_off), 256 entries per block.
template <class FC>
struct EqFamJob {
  const typename FC::T* v;
  typename FC::T* fam_l;
  typename FC::T* fam_r;
  uint32_t h_l, h_r;
  uint32_t block_off;
  uint32_t pad;
};

template <class FC>
__global__ __launch_bounds__(256) void eq_family_batch_kernel(const EqFamJob<FC>* jobs,
                                                              uint32_t njobs) {
  uint32_t lo = 0, hi = njobs;  // last job with block_off <= blockIdx.x
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (jobs[mid].block_off <= blockIdx.x) lo = mid; else hi = mid;
  }
  const EqFamJob<FC> j = jobs[lo];
  __shared__ typename FC::T vs[EQ_STAGE_MAX];
  const typename FC::T* w = eq_stage<FC>(j.v, j.h_l + j.h_r, vs);
  eq_family_entry<FC>(w, j.h_l, j.fam_l, j.h_r, j.fam_r,
                      (size_t)(blockIdx.x - j.block_off) * 256 + threadIdx.x);
}

}  // namespace zs
