// Fiat-Shamir hop run by one full wavefront (64 lanes) on the device.
//
// The SHA-256 chain of a transcript is strictly sequential, so one wave runs
// it. A single wave issues one instruction every four cycles, so the hop's
// time is its instruction count; it is cut two ways:
//  * Every block whose bytes are known when the hop starts (the blocks that
//    absorb the framed message, and the padded tail block(s) of the squeeze,
//    which hold the new pending bytes, the counter and the length) has its
//    message schedule expanded first, lane b expanding block b, into
//    W_t + K_t words in LDS. The chain of compressions then runs the 64
//    rounds alone, reading four words per LDS load. Only the block that
//    absorbs the squeezed bytes (when the pending bytes overflow) depends on
//    the digest; its schedule is expanded when the digest is known.
//  * The chaining values stay in vector registers (vreg): left to itself the
//    compiler runs wave-uniform arithmetic on the scalar unit, which lacks
//    rotates and three-input operations.
// All compressions go through one call site, so the unrolled 64-round body is
// fetched into the instruction cache once per hop.
//
// Byte stream identical to sha256_fs_hop (sha256.h): observe(msg) under slice
// framing, sample(limbs) (reference zorch/sha256_field_transcript.py:100-114,
// 189-196, 251-259; zorch/challenge.py:97-106).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sha256.h"

namespace zs {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct WaveSha {
  uint32_t h[8];
  uint32_t plen;
  uint64_t total;
};

// Keep a wave-uniform value in a vector register. Without this the compiler
// runs uniform 32-bit arithmetic on the scalar unit, which lacks rotates and
// 64-bit multiplies and shuttles every intermediate through readfirstlane — a
// SHA-256 compression then takes ~3x longer than on the vector ALU.
__device__ __forceinline__ uint32_t vreg(uint32_t x) {
  uint32_t y;
  asm("v_mov_b32 %0, %1" : "=v"(y) : "v"(x));
  return y;
}

// Blocks of one hop: at most two absorbing the framed message (pending < 64
// bytes plus at most 20 + 96 framed bytes), two squeeze tail blocks and one
// absorbing the squeezed bytes.
constexpr uint32_t HOP_MAX_BLOCKS = 5;
constexpr uint32_t HOP_MAX_MSG = 96;

// Lanes b < nblk expand block first + b (64 bytes at blocks + 64 (first + b))
// into wk[64 (first + b) + t] = W_t + K_t, t = 0..63. The expansion runs as a
// loop of sixteen-step bodies (the sixteen live words stay in fixed
// registers), which keeps the hop's code small enough to stay in the
// instruction cache beside the other waves' code.
__device__ __forceinline__ void ws_schedule(const uint8_t* blocks, uint32_t first, uint32_t nblk,
                                            uint32_t* wk) {
  const uint32_t lane = threadIdx.x & 63;
  if (lane < nblk) {
    const uint32_t b = first + lane;
    const uint4* b128 = reinterpret_cast<const uint4*>(blocks + 64u * b);
    uint4* o4 = reinterpret_cast<uint4*>(wk + 64u * b);
    uint32_t w[16];
#pragma unroll
    for (int v = 0; v < 4; ++v) {  // all loads before any store
      const uint4 x = b128[v];
      w[4 * v] = __builtin_bswap32(x.x);
      w[4 * v + 1] = __builtin_bswap32(x.y);
      w[4 * v + 2] = __builtin_bswap32(x.z);
      w[4 * v + 3] = __builtin_bswap32(x.w);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v)
      o4[v] = make_uint4(w[4 * v] + ZS_K256[4 * v], w[4 * v + 1] + ZS_K256[4 * v + 1],
                         w[4 * v + 2] + ZS_K256[4 * v + 2], w[4 * v + 3] + ZS_K256[4 * v + 3]);
#pragma unroll 1
    for (int t0 = 16; t0 < 64; t0 += 16) {
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const uint32_t x = w[(u + 1) & 15], y = w[(u + 14) & 15];
        const uint32_t s0 = sha_xor3(sha_rotr(x, 7), sha_rotr(x, 18), x >> 3);
        const uint32_t s1 = sha_xor3(sha_rotr(y, 17), sha_rotr(y, 19), y >> 10);
        w[u] = w[u] + s0 + w[(u + 9) & 15] + s1;
      }
#pragma unroll
      for (int v = 0; v < 4; ++v)
        o4[(t0 >> 2) + v] =
            make_uint4(w[4 * v] + ZS_K256[t0 + 4 * v], w[4 * v + 1] + ZS_K256[t0 + 4 * v + 1],
                       w[4 * v + 2] + ZS_K256[t0 + 4 * v + 2], w[4 * v + 3] + ZS_K256[t0 + 4 * v + 3]);
    }
  }
  wave_sync();
}

// The 64 rounds of one compression from an expanded schedule (LDS), fully
// unrolled (the W + K words read four at a time ahead of their rounds).
__device__ __forceinline__ void sha256_rounds(uint32_t h[8], const uint32_t* wk) {
  uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
  const uint4* q = reinterpret_cast<const uint4*>(wk);
#pragma unroll
  for (int t4 = 0; t4 < 16; ++t4) {
    const uint4 v = q[t4];
    const uint32_t kw[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t S1 = sha_xor3(sha_rotr(e, 6), sha_rotr(e, 11), sha_rotr(e, 25));
      const uint32_t ch = __builtin_amdgcn_bitop3_b32(e, f, g, 0xCA);  // e ? f : g
      const uint32_t t1 = hh + S1 + ch + kw[u];
      const uint32_t S0 = sha_xor3(sha_rotr(a, 2), sha_rotr(a, 13), sha_rotr(a, 22));
      const uint32_t mj = __builtin_amdgcn_bitop3_b32(a, b, c, 0xE8);  // majority
      hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + S0 + mj;
    }
  }
  h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
}

// observe(msg) + sample(limbs), with limbs * item <= 32 (one squeeze counter).
// `msg` (LDS, msg_len <= HOP_MAX_MSG bytes) is framed into `buf` (LDS,
// >= msg_len + 20 bytes); `blocks` (LDS, HOP_MAX_BLOCKS x 64 B, 16-byte
// aligned) and `wk` (LDS, HOP_MAX_BLOCKS x 64 words, 16-byte aligned) hold the hop's blocks and
// schedules, `dgw` (LDS, 8 words) the digest; the squeezed bytes land in
// `out` (LDS). The stream's pending bytes live in `pend` (LDS, 64 B).
__device__ __forceinline__ void ws_fs_hop(WaveSha& s, uint8_t* pend, uint8_t* buf,
                                          const uint8_t* msg, uint32_t msg_len, uint32_t item,
                                          uint32_t limbs, uint8_t* blocks, uint32_t* wk,
                                          uint32_t* dgw, uint8_t* out) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t count = msg_len / item;
  const uint32_t framed = msg_len + 20u;
  // Absorb stream = pending || framed: `na` full blocks, then `rem` bytes
  // pending; the squeeze tail block(s) start with those rem bytes.
  const uint32_t nbytes = item * limbs;
  const uint32_t plen = s.plen;
  const uint32_t total = plen + framed;
  const uint32_t na = total / 64u, rem = total % 64u;
  const uint64_t absorbed = s.total + framed;
  const uint32_t ntail = (rem + 8u + 9u <= 64u) ? 1u : 2u;
  const uint32_t nfinal = (rem + nbytes >= 64u) ? 1u : 0u;
  const uint64_t bits = (absorbed + 8u) * 8u;  // stream || le64(ctr = 0)
  const uint32_t tail_end = 64u * (na + ntail);
  // The framed message, observe(msg) then the squeeze request:
  //   [OP_OBSERVE, KIND_SLICE] le64(count) msg [OP_SQUEEZE, KIND_SLICE] le64(limbs).
  for (uint32_t pos = lane; pos < framed; pos += 64u) {
    uint32_t b;
    if (pos == 0) b = OP_OBSERVE;
    else if (pos == 1) b = KIND_SLICE;
    else if (pos < 10u) b = (uint32_t)(count >> (8u * (pos - 2u))) & 0xffu;
    else if (pos < 10u + msg_len) b = msg[pos - 10u];
    else if (pos == 10u + msg_len) b = OP_SQUEEZE;
    else if (pos == 11u + msg_len) b = KIND_SLICE;
    else b = (uint32_t)((uint64_t)limbs >> (8u * (pos - 12u - msg_len))) & 0xffu;
    buf[pos] = (uint8_t)b;
  }
  wave_sync();
  // The blocks: pending || framed, then the squeeze tail (le64(0) counter,
  // 0x80, zeros, the bit length).
  for (uint32_t pos = lane; pos < tail_end; pos += 64u) {
    uint32_t b;
    if (pos < total) {
      b = pos < plen ? pend[pos] : buf[pos - plen];
    } else {
      const uint32_t p = pos - 64u * na;  // position in the tail
      if (p < rem + 8u) b = 0;            // le64(0)
      else if (p == rem + 8u) b = 0x80u;
      else if (p >= 64u * ntail - 8u) b = (uint32_t)(bits >> (8u * (64u * ntail - 1u - p))) & 0xffu;
      else b = 0;
    }
    blocks[pos] = (uint8_t)b;
  }
  wave_sync();
  ws_schedule(blocks, 0, na + ntail, wk);
  const uint32_t njobs = na + ntail + nfinal;
  uint32_t h[8], d[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    h[i] = s.h[i];
    d[i] = 0;
  }
  for (uint32_t j = 0; j < njobs; ++j) {
    const bool tail = j >= na && j < na + ntail;
    if (j == na + ntail) {
      // Absorb the squeezed bytes: pending (rem bytes, the head of the first
      // tail block) || out[0 .. 64 - rem) fills one more block.
      blocks[64u * j + lane] = lane < rem ? blocks[64u * na + lane] : out[lane - rem];
      wave_sync();
      ws_schedule(blocks, j, 1, wk);
    }
    // The one compression site.
    uint32_t hh[8];
    const bool from_d = tail && j > na;
#pragma unroll
    for (int i = 0; i < 8; ++i) hh[i] = vreg(from_d ? d[i] : h[i]);
    sha256_rounds(hh, wk + 64u * j);
    if (tail) {
#pragma unroll
      for (int i = 0; i < 8; ++i) d[i] = hh[i];
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) h[i] = hh[i];
    }
    if (j == na + ntail - 1u) {
      // Digest complete: its first nbytes (big-endian words) are the output.
      if (lane == 0) {
#pragma unroll
        for (int i = 0; i < 8; ++i) dgw[i] = d[i];
      }
      wave_sync();
      if (lane < nbytes) out[lane] = (uint8_t)(dgw[lane >> 2] >> (24u - 8u * (lane & 3u)));
      wave_sync();
    }
  }
  // New pending bytes: rem || out, minus the block absorbed when it filled.
  uint8_t pb = 0;
  const uint32_t plen2 = rem + nbytes - 64u * nfinal;
  if (lane < plen2) {
    const uint32_t p = lane + 64u * nfinal;  // position in rem || out
    pb = p < rem ? blocks[64u * na + p] : out[p - rem];
  }
  wave_sync();
  pend[lane] = pb;
  wave_sync();
#pragma unroll
  for (int i = 0; i < 8; ++i) s.h[i] = h[i];
  s.plen = plen2;
  s.total = absorbed + nbytes;
}

}  // namespace zs
