// Streaming SHA-256 shared by the host transcript and the device Fiat-Shamir hop.
//
// Replaces hash-frx's ``Sha256State`` stream (``sha256_stream_init / absorb /
// finalize``; reference call sites zorch/sha256_field_transcript.py:45-50,
// 100-114), which is absent from this environment. The state is the FIPS 180-4
// midstate ``h``, the not-yet-compressed tail ``pending`` and the byte counts.
// Squeezes finalize a *view* of the state extended by an 8-byte counter, so
// nothing is copied on the device: the padded tail block(s) are assembled one
// word at a time from ``pending``, the suffix and the length.
#pragma once
#include <stdint.h>
#include <stddef.h>

#include "field.h"

namespace zs {

struct Sha256State {
  uint32_t h[8];
  uint8_t pending[64];
  uint64_t pending_len;
  uint64_t total_len;
};
static_assert(sizeof(Sha256State) == 112, "transcript state layout");

#ifdef __HIP_DEVICE_COMPILE__
#define ZS_K256 zs_k256_dev
#else
#define ZS_K256 zs_k256_host
#endif

#define ZS_K256_VALUES                                                                       \
  {0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, \
   0xab1c5ed5u, 0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, \
   0x9bdc06a7u, 0xc19bf174u, 0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, \
   0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau, 0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, \
   0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u, 0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, \
   0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u, 0xa2bfe8a1u, 0xa81a664bu, \
   0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u, 0x19a4c116u, \
   0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u, \
   0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u, \
   0xc67178f2u}

__device__ __constant__ static const uint32_t zs_k256_dev[64] = ZS_K256_VALUES;
static const uint32_t zs_k256_host[64] = ZS_K256_VALUES;

ZS_HD uint32_t sha_rotr(uint32_t x, int n) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_alignbit(x, x, n);
#else
  return (x >> n) | (x << (32 - n));
#endif
}

// Three-way XOR: one v_bitop3 (truth table 0x96) on the device.
ZS_HD uint32_t sha_xor3(uint32_t a, uint32_t b, uint32_t c) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

// One compression of the 16 big-endian message words w (consumed).
ZS_HD void sha256_compress_words(uint32_t h[8], uint32_t w[16]) {
  uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
#pragma unroll
  for (int t = 0; t < 64; ++t) {
    uint32_t wt;
    if (t < 16) {
      wt = w[t];
    } else {
      uint32_t x = w[(t - 15) & 15], y = w[(t - 2) & 15];
      uint32_t s0 = sha_xor3(sha_rotr(x, 7), sha_rotr(x, 18), x >> 3);
      uint32_t s1 = sha_xor3(sha_rotr(y, 17), sha_rotr(y, 19), y >> 10);
      wt = w[t & 15] + s0 + w[(t - 7) & 15] + s1;
      w[t & 15] = wt;
    }
    uint32_t S1 = sha_xor3(sha_rotr(e, 6), sha_rotr(e, 11), sha_rotr(e, 25));
    uint32_t ch = (e & f) ^ (~e & g);
    uint32_t t1 = hh + S1 + ch + ZS_K256[t] + wt;
    uint32_t S0 = sha_xor3(sha_rotr(a, 2), sha_rotr(a, 13), sha_rotr(a, 22));
    uint32_t mj = (a & b) ^ (c & (a ^ b));
    hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + S0 + mj;
  }
  h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
}

ZS_HD uint32_t sha_be32(const uint8_t* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

ZS_HD void sha256_compress_bytes(uint32_t h[8], const uint8_t* block) {
  uint32_t w[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) w[i] = sha_be32(block + 4 * i);
  sha256_compress_words(h, w);
}

ZS_HD void sha256_init(Sha256State& s) {
  s.h[0] = 0x6a09e667u; s.h[1] = 0xbb67ae85u; s.h[2] = 0x3c6ef372u; s.h[3] = 0xa54ff53au;
  s.h[4] = 0x510e527fu; s.h[5] = 0x9b05688cu; s.h[6] = 0x1f83d9abu; s.h[7] = 0x5be0cd19u;
  for (int i = 0; i < 64; ++i) s.pending[i] = 0;
  s.pending_len = 0;
  s.total_len = 0;
}

ZS_HD void sha256_absorb_byte(Sha256State& s, uint8_t byte) {
  s.pending[s.pending_len] = byte;
  if (++s.pending_len == 64) {
    sha256_compress_bytes(s.h, s.pending);
    s.pending_len = 0;
  }
  s.total_len += 1;
}

ZS_HD void sha256_absorb(Sha256State& s, const uint8_t* data, size_t len) {
  for (size_t i = 0; i < len; ++i) sha256_absorb_byte(s, data[i]);
}

ZS_HD void sha256_absorb_le64(Sha256State& s, uint64_t v) {
  for (int i = 0; i < 8; ++i) sha256_absorb_byte(s, (uint8_t)(v >> (8 * i)));
}

// Digest of the stream extended by `suffix` (slen <= 55 bytes), without
// touching the state: SHA256(stream || suffix).
ZS_HD void sha256_digest_suffix(const Sha256State& s, const uint8_t* suffix, int slen,
                                uint8_t out[32]) {
  uint32_t h[8];
  for (int i = 0; i < 8; ++i) h[i] = s.h[i];
  const int plen = (int)s.pending_len;
  const int mlen = plen + slen;                 // message bytes in the tail
  const int nblk = (mlen + 9 <= 64) ? 1 : 2;    // 0x80 + 8 length bytes
  const uint64_t bits = (s.total_len + (uint64_t)slen) * 8u;
  for (int blk = 0; blk < nblk; ++blk) {
    uint32_t w[16];
    for (int i = 0; i < 16; ++i) {
      uint32_t word = 0;
      for (int j = 0; j < 4; ++j) {
        int pos = blk * 64 + 4 * i + j;
        uint32_t b;
        if (pos < plen) b = s.pending[pos];
        else if (pos < mlen) b = suffix[pos - plen];
        else if (pos == mlen) b = 0x80u;
        else if (pos >= nblk * 64 - 8) b = (uint32_t)(bits >> (8 * (nblk * 64 - 1 - pos))) & 0xffu;
        else b = 0;
        word = (word << 8) | b;
      }
      w[i] = word;
    }
    sha256_compress_words(h, w);
  }
  for (int i = 0; i < 8; ++i) {
    out[4 * i] = (uint8_t)(h[i] >> 24);
    out[4 * i + 1] = (uint8_t)(h[i] >> 16);
    out[4 * i + 2] = (uint8_t)(h[i] >> 8);
    out[4 * i + 3] = (uint8_t)h[i];
  }
}

// Wire vocabulary of the byte transcript (reference zorch/byte_transcript.py:37-43).
constexpr uint8_t OP_DOMAIN = 0x01, OP_LABEL = 0x02, OP_OBSERVE = 0x03, OP_SQUEEZE = 0x04,
                  OP_BYTES = 0x05, KIND_SCALAR = 0x01, KIND_SLICE = 0x02;

// Counter squeeze of nbytes from the current stream, then re-absorb them
// (reference _squeeze_hop, zorch/sha256_field_transcript.py:100-114, without
// the framing absorb, which callers put on the stream first).
ZS_HD void sha256_squeeze(Sha256State& s, uint8_t* out, size_t nbytes) {
  size_t got = 0;
  for (uint64_t ctr = 0; got < nbytes; ++ctr) {
    uint8_t c8[8];
    for (int i = 0; i < 8; ++i) c8[i] = (uint8_t)(ctr >> (8 * i));
    uint8_t dg[32];
    sha256_digest_suffix(s, c8, 8, dg);
    size_t take = nbytes - got < 32 ? nbytes - got : 32;
    for (size_t i = 0; i < take; ++i) out[got + i] = dg[i];
    got += take;
  }
  sha256_absorb(s, out, nbytes);
}

// One Fiat-Shamir hop of the field transcript: observe(values) under slice
// framing, then sample(limbs) (zorch/sha256_field_transcript.py:189-196,
// 251-259; zorch/challenge.py:97-106). `item` is the transcript word size.
ZS_HD void sha256_fs_hop(Sha256State& s, const uint8_t* msg, size_t msg_len, size_t item,
                         size_t limbs, uint8_t* out) {
  sha256_absorb_byte(s, OP_OBSERVE);
  sha256_absorb_byte(s, KIND_SLICE);
  sha256_absorb_le64(s, msg_len / item);
  sha256_absorb(s, msg, msg_len);
  sha256_absorb_byte(s, OP_SQUEEZE);
  sha256_absorb_byte(s, KIND_SLICE);
  sha256_absorb_le64(s, limbs);
  sha256_squeeze(s, out, limbs * item);
}

}  // namespace zs
