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

struct ShaH {
  uint32_t v[8];
};
struct ShaW {
  uint32_t v[16];
};

// One compression, kept out of line so each kernel carries a single copy of
// the fully unrolled 64-round body; state and words travel in registers.
__device__ __noinline__ ShaH ws_compress(ShaH h, ShaW w) {
  sha256_compress_words(h.v, w.v);
  return h;
}

__device__ __forceinline__ void ws_compress_lds(uint32_t h[8], const uint8_t* block) {
  ShaW w;
  const uint32_t* b32 = reinterpret_cast<const uint32_t*>(block);
#pragma unroll
  for (int i = 0; i < 16; ++i) w.v[i] = __builtin_bswap32(b32[i]);
  ShaH s;
#pragma unroll
  for (int i = 0; i < 8; ++i) s.v[i] = h[i];
  s = ws_compress(s, w);
#pragma unroll
  for (int i = 0; i < 8; ++i) h[i] = s.v[i];
}

// Load the wave state from a Sha256State (global or LDS) into registers and
// the pending bytes into the LDS block `pend` (64 B, 4-byte aligned).
__device__ __forceinline__ void ws_load(WaveSha& s, const Sha256State& st, uint8_t* pend) {
  const uint32_t lane = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < 8; ++i) s.h[i] = st.h[i];
  s.plen = (uint32_t)st.pending_len;
  s.total = st.total_len;
  pend[lane] = st.pending[lane];
  wave_sync();
}

__device__ __forceinline__ void ws_store(const WaveSha& s, Sha256State& st, const uint8_t* pend) {
  const uint32_t lane = threadIdx.x & 63;
  st.pending[lane] = lane < s.plen ? pend[lane] : 0;
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) st.h[i] = s.h[i];
    st.pending_len = s.plen;
    st.total_len = s.total;
  }
}

// Absorb `len` bytes of `src` (LDS).
__device__ __forceinline__ void ws_absorb(WaveSha& s, uint8_t* pend, const uint8_t* src, uint32_t len) {
  const uint32_t lane = threadIdx.x & 63;
  uint32_t pos = 0;
  while (pos < len) {
    const uint32_t room = 64u - s.plen;
    const uint32_t take = (len - pos) < room ? (len - pos) : room;
    if (lane < take) pend[s.plen + lane] = src[pos + lane];
    wave_sync();
    s.plen += take;
    s.total += take;
    pos += take;
    if (s.plen == 64u) {
      ws_compress_lds(s.h, pend);
      s.plen = 0;
      wave_sync();
    }
  }
}

// Digest words of SHA256(stream || le64(ctr)); `tail` is 128 B of LDS.
__device__ __forceinline__ void ws_digest_ctr(const WaveSha& s, const uint8_t* pend, uint64_t ctr, uint8_t* tail,
                              uint32_t out[8]) {
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
#pragma unroll
  for (int i = 0; i < 8; ++i) out[i] = s.h[i];
  ws_compress_lds(out, tail);
  if (nblk == 2u) ws_compress_lds(out, tail + 64);
  wave_sync();
}

__device__ __forceinline__ uint32_t ws_pick(const uint32_t w[8], uint32_t i) {
  uint32_t r = w[0];
#pragma unroll
  for (uint32_t k = 1; k < 8; ++k) r = (i == k) ? w[k] : r;
  return r;
}

// observe(msg) + sample(limbs): `msg` (LDS, msg_len bytes) is framed into
// `buf` (LDS, >= msg_len + 20 bytes); the squeezed bytes land in `out` (LDS).
__device__ __forceinline__ void ws_fs_hop(WaveSha& s, uint8_t* pend, uint8_t* buf, const uint8_t* msg,
                          uint32_t msg_len, uint32_t item, uint32_t limbs, uint8_t* tail,
                          uint8_t* out) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t count = msg_len / item;
  const uint32_t total = msg_len + 20u;
  for (uint32_t pos = lane; pos < total; pos += 64u) {
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
  ws_absorb(s, pend, buf, total);
  const uint32_t nbytes = item * limbs;
  for (uint32_t ctr = 0; ctr * 32u < nbytes; ++ctr) {
    uint32_t dg[8];
    ws_digest_ctr(s, pend, ctr, tail, dg);
    const uint32_t take = (nbytes - ctr * 32u) < 32u ? (nbytes - ctr * 32u) : 32u;
    // Digest words to LDS big-endian (static indices), then one byte per lane.
    uint32_t* tw = reinterpret_cast<uint32_t*>(tail);
    if (lane == 0) {
#pragma unroll
      for (int i = 0; i < 8; ++i) tw[i] = __builtin_bswap32(dg[i]);
    }
    wave_sync();
    if (lane < take) out[ctr * 32u + lane] = tail[lane];
    wave_sync();
  }
  ws_absorb(s, pend, out, nbytes);
}

}  // namespace zs
