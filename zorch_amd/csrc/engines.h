// Round-cost engines of the reference beside its linear-time prover: the
// sqrt-space first phase (zorch/sumcheck/sqrt_space.py), the univariate-skip
// round 0 (zorch/sumcheck/univariate_skip.py) and the small-value accumulators
// (zorch/sumcheck/eq/small_value.py, accumulators.py). Each replaces the first
// rounds of a stage; the rounds after them are the dense kernels of kernels.h
// on challenge-field tables the engine leaves in HBM.
#pragma once
#include "kernels.h"

namespace zs {

// ---------------------------------------------------- sqrt-space refold rounds
// Round i < h (h = n/2) of the sqrt-space engine (sqrt_space.py:40-96). The
// input tables are never folded: pair j of round i's table (length
// L = 2^(n-i)) is refolded from them over the bound prefix,
//   lo_q = sum_b E[b] P_q[b L + j],   hi_q = sum_b E[b] P_q[b L + L/2 + j],
// with E = eq(r[<i], .) (2^i challenge-field entries, the newest challenge on
// the index LSB). Round 0 (E_0 = [1]) is the dense round 0 and round 1 forms
// E_1 = [1 - r_0, r_0] from the challenge; later rounds read E from one of two
// buffers, which block 0 extends by the round's challenge once the round is
// finished. A block takes groups of P consecutive pairs
// (coalesced) split over S = 256/P slices of the prefix: while pairs are many
// P = 256 and a thread refolds whole pairs; as they grow few the prefix sum is
// spread over the block and the slices meet in LDS. Slice-0 threads form the
// node products of their pairs; block 0 finishes the round (as in
// pass_kernel) and then extends E by the new challenge for the next round.
// The last refold round (write_work) stores its refolded tables, the round
// table the first dense round folds: the only folded state the engine holds
// is 2 * 2^(n-h) entries a table.
template <class FT, class FC, int K>
struct RefoldLds {
  PassLds<FC, K + 1> ps;
  typename FC::T part[PASS_THREADS][2 * K];
};

template <class FT, class FC, int K>
__global__ __launch_bounds__(PASS_THREADS) void refold_pass_kernel(StageArgs<FT, FC> a,
                                                                   uint32_t round,
                                                                   uint32_t pshift,
                                                                   uint32_t write_work,
                                                                   uint32_t need_claim) {
  using T = typename FC::T;
  constexpr int NA = K + 1;
  constexpr int NPTS = K + 1;
  __shared__ RefoldLds<FT, FC, K> ls;
  RoundLds<FC, NA>& sh = ls.ps.round;
  const uint32_t tid = threadIdx.x, nblk = gridDim.x - 1;
  Ctl<FC>* ctl = a.ctl;
  const size_t half_buf = (size_t)1 << a.sq_half;
  const T* E = a.sq_eq + ((round & 1u) ? half_buf : 0);
  T* E_next = a.sq_eq + ((round & 1u) ? 0 : half_buf);
  // E_1 = [1 - r_0, r_0], from the challenge of round 0 (read before this
  // round's finisher replaces it).
  const T r_prev = ctl->r;
  const T e1_lo = FC::sub(ctl->one, r_prev);
  auto e_at = [&](uint32_t b) -> T { return round == 1 ? (b ? r_prev : e1_lo) : E[b]; };
  if (blockIdx.x == 0) {
    if (tid < 64) {
      fin_load(sh.sm, ctl);
      const uint32_t* from = reinterpret_cast<const uint32_t*>(static_cast<HopLds*>(&sh.sm));
      uint32_t* to = reinterpret_cast<uint32_t*>(&ls.ps.warm);
      for (uint32_t i = tid; i < sizeof(HopLds) / 4; i += 64u) to[i] = from[i];
      wave_sync();
      fs_hop<HOP_PASS>((ZS_LDS HopLds*)&ls.ps.warm, (const ZS_LDS uint8_t*)sh.sm.msg,
                       (uint32_t)(NPTS * sizeof(T)), ctl->item_bytes, ctl->limbs, ctl->tr_kind);
    }
    if (tid == 0) sh.is_last = chain_wait(&ctl->counter, nblk, &ctl->fault) ? 1u : 0u;
    __syncthreads();
    if (!sh.is_last) return;
    slot_rows_sum<FC, NA>(a.slots, sh);
    if (tid < 64) {
      finish_round<FT, FC, K, false, HOP_PASS>(a, round, need_claim != 0, sh.out, sh.sm, a.work,
                                               a.in, a.in_stride, true);
      fin_store(sh.sm, ctl);
    }
    if (tid == 0) __hip_atomic_store(&ctl->counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (round + 1 < a.sq_half) {
      // E_{i+1}[2b + c] = E_i[b] (1 - r) or E_i[b] r (eq.py:55-62, LSB append).
      const T r = sh.sm.r;
      const uint32_t B = 1u << round;
      for (uint32_t x = tid; x < 2 * B; x += PASS_THREADS) {
        const T e = e_at(x >> 1);
        const T hi = FC::mul(e, r);
        E_next[x] = (x & 1u) ? hi : FC::sub(e, hi);
      }
    }
    return;
  }
  const uint32_t blk = blockIdx.x - 1;
  const uint32_t L = 1u << (a.n - round), half = L >> 1;
  const uint32_t P = 1u << pshift, S = PASS_THREADS >> pshift;
  const uint32_t pj = tid & (P - 1), slice = tid >> pshift;
  const uint32_t B = 1u << round;
  const uint32_t groups = half >> pshift;
  T acc[NA];
#pragma unroll
  for (int t = 0; t < NA; ++t) acc[t] = FC::zero();
  for (uint32_t g = blk; g < groups; g += nblk) {
    const uint32_t j = (g << pshift) + pj;
    WeightedSum<FC, FT> lo[K], hi[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      lo[q].init();
      hi[q].init();
    }
    // Two prefix rows a step: the lazy sums take two products between folds.
    uint32_t b = slice;
    for (; b + S < B; b += 2 * S) {
      const T e0 = e_at(b), e1 = e_at(b + S);
      const size_t base0 = (size_t)b * L + j, base1 = (size_t)(b + S) * L + j;
      typename FT::T x[K][4];
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const typename FT::T* tq = a.in + (size_t)q * a.in_stride;
        x[q][0] = tq[base0];
        x[q][1] = tq[base0 + half];
        x[q][2] = tq[base1];
        x[q][3] = tq[base1 + half];
      }
#pragma unroll
      for (int q = 0; q < K; ++q) {
        lo[q].add(e0, x[q][0], false);
        hi[q].add(e0, x[q][1], false);
        lo[q].add(e1, x[q][2]);
        hi[q].add(e1, x[q][3]);
      }
    }
    if (b < B) {
      const T e = e_at(b);
      const size_t base = (size_t)b * L + j;
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const typename FT::T* tq = a.in + (size_t)q * a.in_stride + base;
        lo[q].add(e, tq[0]);
        hi[q].add(e, tq[half]);
      }
    }
    T vlo[K], vhi[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      vlo[q] = lo[q].get();
      vhi[q] = hi[q].get();
    }
    if (S > 1) {
#pragma unroll
      for (int q = 0; q < K; ++q) {
        ls.part[tid][2 * q] = vlo[q];
        ls.part[tid][2 * q + 1] = vhi[q];
      }
      __syncthreads();
      if (slice == 0) {
        for (uint32_t s = 1; s < S; ++s) {
#pragma unroll
          for (int q = 0; q < K; ++q) {
            vlo[q] = FC::add(vlo[q], ls.part[pj + (s << pshift)][2 * q]);
            vhi[q] = FC::add(vhi[q], ls.part[pj + (s << pshift)][2 * q + 1]);
          }
        }
      }
      __syncthreads();
    }
    if (slice == 0) {
      if (write_work) {
#pragma unroll
        for (int q = 0; q < K; ++q) {
          a.work[(size_t)q * a.work_stride + j] = vlo[q];
          a.work[(size_t)q * a.work_stride + j + half] = vhi[q];
        }
      }
      T d[K], v[K];
#pragma unroll
      for (int q = 0; q < K; ++q) {
        d[q] = FC::sub(vhi[q], vlo[q]);
        v[q] = vlo[q];
      }
#pragma unroll
      for (int t = 0; t < NA; ++t) {
        T prod = v[0];
#pragma unroll
        for (int q = 1; q < K; ++q) prod = FC::mul(prod, v[q]);
        acc[t] = FC::add(acc[t], prod);
#pragma unroll
        for (int q = 0; q < K; ++q) v[q] = FC::add(v[q], d[q]);
      }
    }
  }
  block_reduce<FC, NA, PASS_THREADS>(acc, sh.lds, sh.out);
  slot_row_add<FC, NA>(a.slots, blk, sh.out);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (tid == 0)
    __hip_atomic_fetch_add(&ctl->counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Slice shift of a refold round: P = 2^pshift consecutive pairs a group, as
// many as keep at least `want_groups` groups (never fewer than 16 pairs, for
// coalescing, nor more than the round's pairs or 256).
inline uint32_t refold_pshift(size_t pairs, size_t want_groups) {
  uint32_t ps = 8;  // P = 256
  while (ps > 4 && (pairs >> ps) < want_groups) --ps;
  while (ps > 0 && ((size_t)1 << ps) > pairs) --ps;
  return ps;
}

}  // namespace zs
