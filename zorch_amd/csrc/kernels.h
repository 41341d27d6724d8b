// Dense sumcheck round kernels for gfx950 (CDNA4, wave64).
//
// One round of the reference's StandardRound / EqPolyRound
// (zorch/sumcheck/prover.py:157-174, zorch/sumcheck/eq/eq_poly.py:114-221) is
// one streaming pass over HBM that
//   1. folds the previous round's tables at the previous challenge,
//      T'[x] = T[x] + r (T[x + L] - T[x])         (domain.py:226-233, MSB bind),
//   2. evaluates this round's polynomial on the natural domain {0..D},
//      s(t) = sum_j w(j) prod_k (lo_k[j] + t (hi_k[j] - lo_k[j]))
//                                               (domain.py:131-137,158-205),
//      with w the factored eq weight eq_l[j_hi] eq_r[j_lo] in eq mode
//      (eq_poly.py:63-97,144-187), never materialised,
// then a one-block finisher reduces the per-block partials, runs the SHA-256
// Fiat-Shamir hop on the device (sha256_fs_hop), reduces the claim by Lagrange
// interpolation (reduce.py:27-52, poly/univariate.py:36-65) and binds the
// challenge (round.py:41-49). The challenge stays in device memory, so a whole
// stage is a chain of launches on one stream with no host round trip.
//
// Layout: tables are table-major (table q at q*stride), element-contiguous;
// BabyBear^4 elements are 16 B and move as one dwordx4. Pair j of a round with
// table length L reads T[j] and T[j + L/2]; a folding pass reads the four
// quarter-strided source entries j + {0, L/2, L, 3L/2} and writes the two
// folded entries j and j + L/2 — only positions thread j itself read, so the
// fold runs in place after the first round. Consecutive lanes touch
// consecutive addresses in all four streams.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "field.h"
#include "sha256.h"

namespace zs {

constexpr int MAXPTS = 6;     // degree <= 4 product + eq factor
constexpr int MAXK = 4;
constexpr int PASS_THREADS = 256;
constexpr int TAIL_THREADS = 1024;
constexpr int MAXG = 4096;    // max blocks of one pass (partials capacity)

// Device control block of one proof (lives in device memory).
template <class FC>
struct Ctl {
  Sha256State sha;
  uint32_t item_bytes;   // transcript word size
  uint32_t limbs;        // transcript words per challenge
  uint32_t counter;
  uint32_t pad;
  typename FC::T claim;
  typename FC::T r;        // current challenge, canonical
  typename FC::T eq_prev;  // eq mass of the bound prefix (eq mode)
  typename FC::T one;
  typename FC::T lag_inv[MAXPTS];  // 1 / prod_{j != t} (t - j) over {0..npts-1}
};

template <class FT, class FC>
struct StageArgs {
  const typename FT::T* in;     // input tables (k x 2^n), table field
  size_t in_stride;
  typename FC::T* work;         // folded tables (k x 2^(n-1)), challenge field
  size_t work_stride;
  uint32_t n;
  uint32_t k;
  Ctl<FC>* ctl;
  typename FC::T* partials;     // MAXG x MAXPTS
  typename FC::T* msgs;         // n x npts
  uint8_t* point;               // n x sizeof(FC::T) raw challenge bytes
  typename FC::T* final_out;    // k (+1) values, may be null
  const typename FC::T* eq_point;  // n coordinates (eq mode)
  const typename FC::T* eq_l_fam;  // suffix family of eq_point[0 : n/2]
  const typename FC::T* eq_r_fam;  // suffix family of eq_point[n/2 : n]
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

// Block-wide field sum of NPTS accumulators; the result lands in out[] (LDS)
// and is valid for every thread after the trailing barrier.
template <class F, int NPTS, int THREADS>
__device__ void block_reduce(typename F::T (&acc)[NPTS], typename F::T* lds,
                             typename F::T* out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < NPTS; ++t) {
    typename F::T v = acc[t];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = F::add(v, shfl_xor_el<F>(v, m));
    acc[t] = v;
  }
  if (lane == 0) {
#pragma unroll
    for (int t = 0; t < NPTS; ++t) lds[wave * NPTS + t] = acc[t];
  }
  __syncthreads();
  if (threadIdx.x < NPTS) {
    typename F::T s = F::zero();
    for (int w = 0; w < THREADS / 64; ++w) s = F::add(s, lds[w * NPTS + threadIdx.x]);
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// ------------------------------------------------------------- eq weights
// Round index i = round + 1 (reference numbering). Returns the weight tables
// for this round: eq_wl (null when only one half weights) and eq_wr, with
// nxr = |eq_wr| (eq_poly.py:144-187).
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
// Accumulate this thread's pairs j = start, start+stride, ... < L/2 of round
// `round`. FOLD: source tables have length 2L and are folded at r into dst
// first; else the source tables are this round's tables (length L).
template <class FS, class FC, int K, int NPTS, bool EQ, bool FOLD>
__device__ __forceinline__ void pass_accumulate(
    const typename FS::T* src, size_t src_stride, typename FC::T* dst,
    size_t dst_stride, uint32_t L, typename FC::T r, const typename FC::T* wl,
    const typename FC::T* wr, uint32_t nxr_log, size_t start, size_t stride,
    typename FC::T (&acc)[NPTS]) {
  // Values are computed in FV: the source field before the first fold, the
  // challenge field after.
  using FV = typename std::conditional<FOLD, FC, FS>::type;
  using V = typename FV::T;
  const size_t half = L >> 1;
  const size_t nxr_mask = ((size_t)1 << nxr_log) - 1;
  for (size_t j = start; j < half; j += stride) {
    V lo[K], df[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const typename FS::T* s = src + (size_t)q * src_stride;
      if constexpr (FOLD) {
        typename FS::T s0 = s[j], s1 = s[j + L], s2 = s[j + half], s3 = s[j + half + L];
        V a = FC::add(Lift<FS, FC>::go(s0), MixMul<FC, FS>::go(r, FS::sub(s1, s0)));
        V b = FC::add(Lift<FS, FC>::go(s2), MixMul<FC, FS>::go(r, FS::sub(s3, s2)));
        typename FC::T* d = dst + (size_t)q * dst_stride;
        d[j] = a;
        d[j + half] = b;
        lo[q] = a;
        df[q] = FC::sub(b, a);
      } else {
        V a = s[j], b = s[j + half];
        lo[q] = a;
        df[q] = FV::sub(b, a);
      }
    }
    typename FC::T w;
    if constexpr (EQ) {
      w = wl ? FC::mul(wl[j >> nxr_log], wr[j & nxr_mask]) : wr[j];
    }
#pragma unroll
    for (int t = 0; t < NPTS; ++t) {
      V prod = lo[0];
#pragma unroll
      for (int q = 1; q < K; ++q) prod = FV::mul(prod, lo[q]);
      if constexpr (EQ) {
        acc[t] = FC::add(acc[t], MixMul<FC, FV>::go(w, prod));
      } else {
        acc[t] = FC::add(acc[t], Lift<FV, FC>::go(prod));
      }
      if (t + 1 < NPTS) {
#pragma unroll
        for (int q = 0; q < K; ++q) lo[q] = FV::add(lo[q], df[q]);
      }
    }
  }
}

// --------------------------------------------------------- element bytes
template <class F>
__device__ __forceinline__ void el_to_bytes(typename F::T v, uint8_t* out);
template <>
__device__ __forceinline__ void el_to_bytes<BB>(uint32_t v, uint8_t* out) {
  for (int i = 0; i < 4; ++i) out[i] = (uint8_t)(v >> (8 * i));
}
template <>
__device__ __forceinline__ void el_to_bytes<BB4>(bb4 v, uint8_t* out) {
  for (int c = 0; c < 4; ++c)
    for (int i = 0; i < 4; ++i) out[4 * c + i] = (uint8_t)(v.c[c] >> (8 * i));
}
template <>
__device__ __forceinline__ void el_to_bytes<GL>(uint64_t v, uint8_t* out) {
  for (int i = 0; i < 8; ++i) out[i] = (uint8_t)(v >> (8 * i));
}
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

// ----------------------------------------------------------- finisher
// Thread 0 only: message from the reduced sums, FS hop, claim reduction,
// bind, eq-mass update, and (last round) the final fold.
template <class FS, class FC, int K, int NPTS, bool EQ>
__device__ void finish_round(const StageArgs<FS, FC>& a, uint32_t round,
                             const typename FC::T* sums, const typename FS::T* last_src,
                             size_t last_src_stride, bool last_src_is_input,
                             Sha256State* sh, uint8_t* bytes) {
  using T = typename FC::T;
  Ctl<FC>* ctl = a.ctl;
  T msg[NPTS];
  T w_i = FC::zero();
  if constexpr (EQ) {
    w_i = a.eq_point[round];
    T eqp = ctl->eq_prev;
    T l1 = FC::mul(eqp, w_i);
    T l0 = FC::sub(eqp, l1);
    T ld = FC::sub(l1, l0);
    T lt = l0;
#pragma unroll
    for (int t = 0; t < NPTS; ++t) {
      msg[t] = FC::mul(lt, sums[t]);
      lt = FC::add(lt, ld);
    }
  } else {
#pragma unroll
    for (int t = 0; t < NPTS; ++t) msg[t] = sums[t];
  }
  constexpr int EB = FC::bytes;
#pragma unroll
  for (int t = 0; t < NPTS; ++t) {
    a.msgs[(size_t)round * NPTS + t] = msg[t];
    el_to_bytes<FC>(msg[t], bytes + t * EB);
  }
  *sh = ctl->sha;
  uint8_t* raw = bytes + NPTS * EB;
  sha256_fs_hop(*sh, bytes, (size_t)NPTS * EB, ctl->item_bytes, ctl->limbs, raw);
  ctl->sha = *sh;
  for (int i = 0; i < EB; ++i) a.point[(size_t)round * EB + i] = raw[i];
  T r = el_from_bytes<FC>(raw);
  // Lagrange basis at r over {0..NPTS-1}: prefix/suffix products of (r - j).
  T diff[NPTS], pre[NPTS + 1], suf[NPTS + 1];
#pragma unroll
  for (int j = 0; j < NPTS; ++j) diff[j] = FC::sub(r, FC::from_int(j));
  pre[0] = FC::one();
  suf[NPTS] = FC::one();
#pragma unroll
  for (int j = 0; j < NPTS; ++j) pre[j + 1] = FC::mul(pre[j], diff[j]);
#pragma unroll
  for (int j = NPTS - 1; j >= 0; --j) suf[j] = FC::mul(suf[j + 1], diff[j]);
  T claim = FC::zero();
#pragma unroll
  for (int t = 0; t < NPTS; ++t)
    claim = FC::add(claim, FC::mul(msg[t], FC::mul(FC::mul(pre[t], suf[t + 1]), ctl->lag_inv[t])));
  ctl->claim = claim;
  ctl->r = r;
  if constexpr (EQ) {
    T one = FC::one();
    T ef = FC::add(FC::mul(r, w_i), FC::mul(FC::sub(one, r), FC::sub(one, w_i)));
    ctl->eq_prev = FC::mul(ctl->eq_prev, ef);
  }
  if (round + 1 == a.n && a.final_out) {
    for (uint32_t q = 0; q < a.k; ++q) {
      T lo, hi;
      if (last_src_is_input) {
        lo = Lift<FS, FC>::go(last_src[(size_t)q * last_src_stride]);
        hi = Lift<FS, FC>::go(last_src[(size_t)q * last_src_stride + 1]);
      } else {
        const T* s = reinterpret_cast<const T*>(last_src);
        lo = s[(size_t)q * last_src_stride];
        hi = s[(size_t)q * last_src_stride + 1];
      }
      a.final_out[q] = FC::add(lo, FC::mul(r, FC::sub(hi, lo)));
    }
    if constexpr (EQ) a.final_out[a.k] = ctl->eq_prev;
  }
}

// ------------------------------------------------------------- kernels
// Multi-block pass of one round: partial sums per block.
template <class FT, class FS, class FC, int K, int NPTS, bool EQ, bool FOLD>
__global__ __launch_bounds__(PASS_THREADS) void pass_kernel(StageArgs<FT, FC> a,
                                                            const typename FS::T* src,
                                                            size_t src_stride, uint32_t L,
                                                            uint32_t round) {
  using T = typename FC::T;
  __shared__ T lds[(PASS_THREADS / 64) * NPTS];
  __shared__ T out[NPTS];
  T acc[NPTS];
#pragma unroll
  for (int t = 0; t < NPTS; ++t) acc[t] = FC::zero();
  const T* wl = nullptr;
  const T* wr = nullptr;
  uint32_t nxr_log = 0;
  if constexpr (EQ) select_eq<FC>(round, a.n, a.eq_l_fam, a.eq_r_fam, &a.ctl->one, wl, wr, nxr_log);
  T r = FC::zero();
  if constexpr (FOLD) r = a.ctl->r;
  const size_t start = (size_t)blockIdx.x * PASS_THREADS + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * PASS_THREADS;
  pass_accumulate<FS, FC, K, NPTS, EQ, FOLD>(src, src_stride, a.work, a.work_stride, L, r, wl,
                                             wr, nxr_log, start, stride, acc);
  block_reduce<FC, NPTS, PASS_THREADS>(acc, lds, out);
  if (threadIdx.x < NPTS) a.partials[(size_t)blockIdx.x * NPTS + threadIdx.x] = out[threadIdx.x];
}

// One-block finisher of a multi-block pass.
template <class FS, class FC, int K, int NPTS, bool EQ>
__global__ __launch_bounds__(PASS_THREADS) void finish_kernel(StageArgs<FS, FC> a, uint32_t nblocks,
                                                              uint32_t round,
                                                              const typename FS::T* last_src,
                                                              size_t last_src_stride,
                                                              int last_src_is_input) {
  using T = typename FC::T;
  __shared__ T lds[(PASS_THREADS / 64) * NPTS];
  __shared__ T out[NPTS];
  __shared__ Sha256State sh;
  __shared__ uint8_t bytes[MAXPTS * 16 + 64];
  T acc[NPTS];
#pragma unroll
  for (int t = 0; t < NPTS; ++t) acc[t] = FC::zero();
  for (uint32_t b = threadIdx.x; b < nblocks; b += PASS_THREADS) {
#pragma unroll
    for (int t = 0; t < NPTS; ++t) acc[t] = FC::add(acc[t], a.partials[(size_t)b * NPTS + t]);
  }
  block_reduce<FC, NPTS, PASS_THREADS>(acc, lds, out);
  if (threadIdx.x == 0)
    finish_round<FS, FC, K, NPTS, EQ>(a, round, out, last_src, last_src_stride,
                                      last_src_is_input != 0, &sh, bytes);
}

// Single-block kernel for the small tail rounds [first, n): each round is a
// block-stride pass, a block reduction and the finisher, separated by
// barriers. Global writes of one round are visible to the block in the next.
template <class FS, class FC, int K, int NPTS, bool EQ>
__global__ __launch_bounds__(TAIL_THREADS) void tail_kernel(StageArgs<FS, FC> a, uint32_t first,
                                                            uint32_t last) {
  using T = typename FC::T;
  __shared__ T lds[(TAIL_THREADS / 64) * NPTS];
  __shared__ T out[NPTS];
  __shared__ Sha256State sh;
  __shared__ uint8_t bytes[MAXPTS * 16 + 64];
  for (uint32_t round = first; round < last; ++round) {
    const uint32_t L = 1u << (a.n - round);
    T acc[NPTS];
#pragma unroll
    for (int t = 0; t < NPTS; ++t) acc[t] = FC::zero();
    const T* wl = nullptr;
    const T* wr = nullptr;
    uint32_t nxr_log = 0;
    if constexpr (EQ) select_eq<FC>(round, a.n, a.eq_l_fam, a.eq_r_fam, &a.ctl->one, wl, wr, nxr_log);
    const T r = a.ctl->r;
    if (round == 0) {
      pass_accumulate<FS, FC, K, NPTS, EQ, false>(a.in, a.in_stride, a.work, a.work_stride, L,
                                                  r, wl, wr, nxr_log, threadIdx.x,
                                                  TAIL_THREADS, acc);
    } else if (round == 1) {
      pass_accumulate<FS, FC, K, NPTS, EQ, true>(a.in, a.in_stride, a.work, a.work_stride, L,
                                                 r, wl, wr, nxr_log, threadIdx.x, TAIL_THREADS,
                                                 acc);
    } else {
      pass_accumulate<FC, FC, K, NPTS, EQ, true>(a.work, a.work_stride, a.work, a.work_stride,
                                                 L, r, wl, wr, nxr_log, threadIdx.x,
                                                 TAIL_THREADS, acc);
    }
    block_reduce<FC, NPTS, TAIL_THREADS>(acc, lds, out);
    if (threadIdx.x == 0) {
      if (round == 0)
        finish_round<FS, FC, K, NPTS, EQ>(a, round, out, a.in, a.in_stride, true, &sh, bytes);
      else
        finish_round<FS, FC, K, NPTS, EQ>(a, round, out,
                                          reinterpret_cast<const typename FS::T*>(a.work),
                                          a.work_stride, false, &sh, bytes);
    }
    __threadfence_block();
    __syncthreads();
  }
}

// Suffix eq family of v[0:h] (compute_eq_evaluations, eq_poly.py:46-51):
// member m = eq(v[h-1-m:], .) with the first coordinate on the MSB, stored at
// fam_offset(m). One block; member m+1 = [member m * (1 - c), member m * c].
template <class FC>
__global__ __launch_bounds__(256) void eq_family_kernel(const typename FC::T* v, uint32_t h,
                                                        typename FC::T* fam) {
  using T = typename FC::T;
  for (uint32_t m = 0; m < h; ++m) {
    const T c = v[h - 1 - m];
    T* dstm = fam + fam_offset(m);
    const size_t len = (size_t)1 << m;  // previous member length
    for (size_t x = threadIdx.x; x < len; x += blockDim.x) {
      T prev = (m == 0) ? FC::one() : fam[fam_offset(m - 1) + x];
      T hi = FC::mul(prev, c);
      dstm[x] = FC::sub(prev, hi);
      dstm[x + len] = hi;
    }
    __threadfence_block();
    __syncthreads();
  }
}

}  // namespace zs
