// Streaming BLAKE3 (hash mode) shared by the host transcript and the device
// Fiat-Shamir hop.
//
// Replaces hash-frx's ``Blake3Stream`` (absorb / finalize; reference call
// sites zorch/blake3_field_transcript.py:206-215, 281-409), which is absent
// from this environment, from the BLAKE3 specification. The state is the
// current chunk (chaining value, the not-yet-compressed block, blocks done,
// chunk index) and the stack of completed subtree chaining values. A full
// block is compressed only once more input arrives, because the last block of
// a chunk carries CHUNK_END (and ROOT when it is the whole tree).
// Finalisation reads extendable output without changing the state.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "field.h"

namespace zs {

constexpr uint32_t B3_CHUNK_START = 1, B3_CHUNK_END = 2, B3_PARENT = 4, B3_ROOT = 8;
constexpr uint32_t B3_BLOCK = 64, B3_CHUNK = 1024;
constexpr uint32_t B3_MAX_DEPTH = 16;  // subtree stack: streams up to 2^16 chunks (64 MiB)

struct Blake3State {
  uint32_t cv[8];                    // chaining value of the current chunk so far
  uint32_t stack[B3_MAX_DEPTH][8];   // chaining values of completed subtrees
  uint8_t block[64];                 // bytes of the current block
  uint64_t chunk_counter;            // index of the current chunk
  uint32_t block_len;                // bytes in `block`
  uint32_t blocks_done;              // blocks of the current chunk already compressed
  uint32_t stack_len;
  uint32_t reserved;
};
static_assert(sizeof(Blake3State) == 632, "transcript state layout");

ZS_HD uint32_t b3_iv(int i) {
  switch (i) {
    case 0: return 0x6A09E667u; case 1: return 0xBB67AE85u;
    case 2: return 0x3C6EF372u; case 3: return 0xA54FF53Au;
    case 4: return 0x510E527Fu; case 5: return 0x9B05688Cu;
    case 6: return 0x1F83D9ABu; default: return 0x5BE0CD19u;
  }
}

ZS_HD uint32_t b3_rotr(uint32_t x, int n) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_alignbit(x, x, n);
#else
  return (x >> n) | (x << (32 - n));
#endif
}

#define ZS_B3_G(a, b, c, d, mx, my)      \
  do {                                   \
    v[a] = v[a] + v[b] + (mx);           \
    v[d] = b3_rotr(v[d] ^ v[a], 16);     \
    v[c] = v[c] + v[d];                  \
    v[b] = b3_rotr(v[b] ^ v[c], 12);     \
    v[a] = v[a] + v[b] + (my);           \
    v[d] = b3_rotr(v[d] ^ v[a], 8);      \
    v[c] = v[c] + v[d];                  \
    v[b] = b3_rotr(v[b] ^ v[c], 7);      \
  } while (0)

// The compression function: out[0..16) (out[0..8) is the new chaining value;
// all 16 words are extended output at the root). `v` in registers, the
// message schedule applied by index (the fixed permutation, composed).
ZS_HD void b3_compress(const uint32_t cv[8], const uint32_t m0[16], uint64_t counter,
                       uint32_t block_len, uint32_t flags, uint32_t out[16]) {
  uint32_t v[16], m[16];
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = cv[i];
#pragma unroll
  for (int i = 0; i < 4; ++i) v[8 + i] = b3_iv(i);
  v[12] = (uint32_t)counter;
  v[13] = (uint32_t)(counter >> 32);
  v[14] = block_len;
  v[15] = flags;
#pragma unroll
  for (int i = 0; i < 16; ++i) m[i] = m0[i];
#pragma unroll
  for (int r = 0; r < 7; ++r) {
    ZS_B3_G(0, 4, 8, 12, m[0], m[1]);
    ZS_B3_G(1, 5, 9, 13, m[2], m[3]);
    ZS_B3_G(2, 6, 10, 14, m[4], m[5]);
    ZS_B3_G(3, 7, 11, 15, m[6], m[7]);
    ZS_B3_G(0, 5, 10, 15, m[8], m[9]);
    ZS_B3_G(1, 6, 11, 12, m[10], m[11]);
    ZS_B3_G(2, 7, 8, 13, m[12], m[13]);
    ZS_B3_G(3, 4, 9, 14, m[14], m[15]);
    if (r < 6) {  // m <- m[MSG_PERMUTATION]
      const uint32_t t[16] = {m[2], m[6], m[3], m[10], m[7], m[0], m[4], m[13],
                              m[1], m[11], m[12], m[5], m[9], m[14], m[15], m[8]};
#pragma unroll
      for (int i = 0; i < 16; ++i) m[i] = t[i];
    }
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    out[i] = v[i] ^ v[i + 8];
    out[i + 8] = v[i + 8] ^ cv[i];
  }
}
#undef ZS_B3_G

ZS_HD void b3_words(const uint8_t* p, uint32_t m[16]) {
  for (int i = 0; i < 16; ++i)
    m[i] = (uint32_t)p[4 * i] | ((uint32_t)p[4 * i + 1] << 8) | ((uint32_t)p[4 * i + 2] << 16) |
           ((uint32_t)p[4 * i + 3] << 24);
}

ZS_HD void b3_init(Blake3State& s) {
  for (int i = 0; i < 8; ++i) s.cv[i] = b3_iv(i);
  for (uint32_t d = 0; d < B3_MAX_DEPTH; ++d)
    for (int i = 0; i < 8; ++i) s.stack[d][i] = 0;
  for (int i = 0; i < 64; ++i) s.block[i] = 0;
  s.chunk_counter = 0;
  s.block_len = 0;
  s.blocks_done = 0;
  s.stack_len = 0;
  s.reserved = 0;
}

// Chaining value of a parent node over two child chaining values.
ZS_HD void b3_parent_cv(const uint32_t left[8], const uint32_t right[8], uint32_t out8[8]) {
  uint32_t m[16], iv[8], o[16];
  for (int i = 0; i < 8; ++i) {
    m[i] = left[i];
    m[8 + i] = right[i];
    iv[i] = b3_iv(i);
  }
  b3_compress(iv, m, 0, B3_BLOCK, B3_PARENT, o);
  for (int i = 0; i < 8; ++i) out8[i] = o[i];
}

// Absorb `len` bytes; returns false if the stream outgrows the subtree stack.
ZS_HD bool b3_absorb(Blake3State& s, const uint8_t* data, size_t len) {
  while (len) {
    if (s.blocks_done * B3_BLOCK + s.block_len == B3_CHUNK) {
      // The chunk is complete and more input follows: push its chaining
      // value, merging completed pairs of subtrees.
      uint32_t m[16], o[16], ncv[8];
      b3_words(s.block, m);
      b3_compress(s.cv, m, s.chunk_counter, s.block_len, B3_CHUNK_END, o);
      for (int i = 0; i < 8; ++i) ncv[i] = o[i];
      uint64_t total = s.chunk_counter + 1;
      while ((total & 1) == 0) {
        s.stack_len -= 1;
        b3_parent_cv(s.stack[s.stack_len], ncv, ncv);
        total >>= 1;
      }
      if (s.stack_len >= B3_MAX_DEPTH) return false;
      for (int i = 0; i < 8; ++i) s.stack[s.stack_len][i] = ncv[i];
      s.stack_len += 1;
      for (int i = 0; i < 8; ++i) s.cv[i] = b3_iv(i);
      s.chunk_counter += 1;
      s.block_len = 0;
      s.blocks_done = 0;
    }
    if (s.block_len == B3_BLOCK) {
      uint32_t m[16], o[16];
      b3_words(s.block, m);
      b3_compress(s.cv, m, s.chunk_counter, B3_BLOCK, s.blocks_done == 0 ? B3_CHUNK_START : 0, o);
      for (int i = 0; i < 8; ++i) s.cv[i] = o[i];
      s.blocks_done += 1;
      s.block_len = 0;
    }
    const size_t room = B3_BLOCK - s.block_len;
    const size_t take = len < room ? len : room;
    for (size_t i = 0; i < take; ++i) s.block[s.block_len + i] = data[i];
    for (size_t i = s.block_len + take; i < B3_BLOCK; ++i) s.block[i] = 0;
    s.block_len += (uint32_t)take;
    data += take;
    len -= take;
  }
  return true;
}

// The root node of the stream (its input chaining value, block words, block
// length and flags, without ROOT): the chunk node when there is one chunk,
// else the parent chain folded over the subtree stack.
ZS_HD void b3_root(const Blake3State& s, uint32_t cv[8], uint32_t m[16], uint32_t* block_len,
                   uint32_t* flags) {
  for (int i = 0; i < 8; ++i) cv[i] = s.cv[i];
  b3_words(s.block, m);
  *block_len = s.block_len;
  *flags = B3_CHUNK_END | (s.blocks_done == 0 ? B3_CHUNK_START : 0);
  uint64_t counter = s.chunk_counter;
  for (uint32_t d = s.stack_len; d-- > 0;) {
    uint32_t o[16];
    b3_compress(cv, m, counter, *block_len, *flags, o);  // child chaining value
    for (int i = 0; i < 8; ++i) {
      m[i] = s.stack[d][i];
      m[8 + i] = o[i];
      cv[i] = b3_iv(i);
    }
    counter = 0;
    *block_len = B3_BLOCK;
    *flags = B3_PARENT;
  }
}

// Extendable output of everything absorbed, nbytes of it; the state is unchanged.
ZS_HD void b3_finalize(const Blake3State& s, uint8_t* out, size_t nbytes) {
  uint32_t cv[8], m[16], block_len, flags;
  b3_root(s, cv, m, &block_len, &flags);
  for (uint64_t blk = 0; blk * 64 < nbytes; ++blk) {
    uint32_t o[16];
    // At the root the counter field numbers the 64-byte output blocks.
    b3_compress(cv, m, blk, block_len, flags | B3_ROOT, o);
    for (int i = 0; i < 16 && blk * 64 + 4 * i < nbytes; ++i)
      for (int j = 0; j < 4 && blk * 64 + 4 * i + j < nbytes; ++j)
        out[blk * 64 + 4 * i + j] = (uint8_t)(o[i] >> (8 * j));
  }
}

// One Fiat-Shamir hop of the BLAKE3 field transcript: observe(values) under
// slice framing, sample(limbs): absorb [OP_SQUEEZE, KIND_SLICE] || le64(limbs),
// read limbs * item bytes of extendable output, re-absorb them
// (zorch/blake3_field_transcript.py:206-215, 320-330, 387-395).
ZS_HD bool b3_fs_hop(Blake3State& s, const uint8_t* msg, size_t msg_len, size_t item,
                     size_t limbs, uint8_t* out) {
  uint8_t frame[10];
  frame[0] = 0x03;  // OP_OBSERVE
  frame[1] = 0x02;  // KIND_SLICE
  const uint64_t count = msg_len / item;
  for (int i = 0; i < 8; ++i) frame[2 + i] = (uint8_t)(count >> (8 * i));
  bool ok = b3_absorb(s, frame, 10) && b3_absorb(s, msg, msg_len);
  frame[0] = 0x04;  // OP_SQUEEZE
  frame[1] = 0x02;
  for (int i = 0; i < 8; ++i) frame[2 + i] = (uint8_t)((uint64_t)limbs >> (8 * i));
  ok = ok && b3_absorb(s, frame, 10);
  const size_t nbytes = limbs * item;
  b3_finalize(s, out, nbytes);
  return ok && b3_absorb(s, out, nbytes);
}

}  // namespace zs
