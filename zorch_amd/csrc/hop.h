// Fiat-Shamir hop run by one full wavefront (64 lanes) on the device.
//
// The SHA-256 chain of a transcript is strictly sequential, so one wave runs
// it; the lanes share the byte work around the compressions: absorbing copies
// up to 64 bytes into the pending block in one step (one byte per lane), the
// padded tail block(s) of a squeeze are assembled one byte per lane, and the
// compression reads its 16 message words from LDS as aligned dwords. The
// chaining value, pending length and total length are held uniformly in every
// lane's registers; only the 64-byte pending block lives in LDS.
//
// A hop is a small schedule of compressions (absorb the framed message, one
// or two tail blocks per squeeze counter, absorb the squeezed bytes). They all
// go through one call site of the fully unrolled compression, so the 64-round
// body is fetched into the instruction cache once per hop and then reused.
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

// Fill the padded tail of SHA256(stream || le64(ctr)) into `tail` (128 B of
// LDS), one byte per lane; returns the number of tail blocks (1 or 2).
__device__ __forceinline__ uint32_t ws_build_tail(const WaveSha& s, const uint8_t* pend,
                                                  uint64_t ctr, uint8_t* tail) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t plen = s.plen, mlen = plen + 8u;
  const uint32_t nblk = (mlen + 9u <= 64u) ? 1u : 2u;
  const uint64_t bits = (s.total + 8u) * 8u;
  for (uint32_t half = 0; half < nblk; ++half) {
    const uint32_t pos = half * 64u + lane;
    uint32_t b;
    if (pos < plen) b = pend[pos];
    else if (pos < mlen) b = (uint32_t)(ctr >> (8u * (pos - plen))) & 0xffu;
    else if (pos == mlen) b = 0x80u;
    else if (pos >= nblk * 64u - 8u) b = (uint32_t)(bits >> (8u * (nblk * 64u - 1u - pos))) & 0xffu;
    else b = 0;
    tail[pos] = (uint8_t)b;
  }
  wave_sync();
  return nblk;
}

// observe(msg) + sample(limbs): `msg` (LDS, msg_len bytes) is framed into
// `buf` (LDS, >= msg_len + 20 bytes); the squeezed bytes land in `out` (LDS,
// 32-byte granules), `tail` (LDS, 128 B) holds squeeze tail blocks and `dgw`
// (LDS, 8 words) the chaining value between the two blocks of a squeeze
// tail, so that only the stream state occupies registers across compressions.
__device__ __forceinline__ void ws_fs_hop(WaveSha& s, uint8_t* pend, uint8_t* buf,
                                          const uint8_t* msg, uint32_t msg_len, uint32_t item,
                                          uint32_t limbs, uint8_t* tail, uint32_t* dgw,
                                          uint8_t* out) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t count = msg_len / item;
  const uint32_t framed = msg_len + 20u;
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
  const uint32_t nbytes = item * limbs;
  // phase 0: absorb buf; 1: squeeze tail blocks; 2: absorb out; 3: done.
  uint32_t phase = 0;
  const uint8_t* src = buf;
  uint32_t src_len = framed, src_pos = 0;
  uint32_t ctr = 0, dblk = 0, dnblk = 0, got = 0;
  while (phase != 3u) {
    const uint8_t* blk;
    bool squeeze;
    if (phase != 1u) {
      const uint32_t room = 64u - s.plen;
      const uint32_t take = (src_len - src_pos) < room ? (src_len - src_pos) : room;
      if (lane < take) pend[s.plen + lane] = src[src_pos + lane];
      wave_sync();
      s.plen += take;
      s.total += take;
      src_pos += take;
      if (s.plen < 64u) {  // source exhausted, block not full
        if (phase == 0u) {
          phase = 1u;
          dblk = 0;
          continue;
        }
        phase = 3u;
        break;
      }
      blk = pend;
      squeeze = false;
    } else {
      if (dblk == 0) dnblk = ws_build_tail(s, pend, ctr, tail);
      blk = tail + 64u * dblk;
      squeeze = true;
    }
    // The one compression site.
    uint32_t w[16], hh[8];
    const uint32_t* b32 = reinterpret_cast<const uint32_t*>(blk);
#pragma unroll
    for (int i = 0; i < 16; ++i) w[i] = vreg(__builtin_bswap32(b32[i]));
    const bool from_dg = squeeze && dblk > 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) hh[i] = vreg(from_dg ? dgw[i] : s.h[i]);
    sha256_compress_words(hh, w);
    wave_sync();
    if (!squeeze) {
#pragma unroll
      for (int i = 0; i < 8; ++i) s.h[i] = hh[i];
      s.plen = 0;
      continue;
    }
    if (lane == 0) {
#pragma unroll
      for (int i = 0; i < 8; ++i) dgw[i] = hh[i];
    }
    wave_sync();
    if (++dblk < dnblk) continue;
    // Digest complete: its first bytes (big-endian words) are squeeze output.
    const uint32_t take = (nbytes - got) < 32u ? (nbytes - got) : 32u;
    if (lane < take)
      out[got + lane] = (uint8_t)(dgw[lane >> 2] >> (24u - 8u * (lane & 3u)));
    wave_sync();
    got += take;
    ++ctr;
    dblk = 0;
    if (got >= nbytes) {
      phase = 2u;
      src = out;
      src_len = nbytes;
      src_pos = 0;
    }
  }
}

}  // namespace zs
