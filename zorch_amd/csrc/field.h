// Prime- and extension-field arithmetic for the dense sumcheck kernels.
//
// Replaces the third-party field dtypes the reference computes in
// (zk_dtypes babybear_mont / babybearx4_mont / goldilocks, used at
// zorch/sumcheck/domain.py:101-102,232-233 and zorch/poly/univariate.py:55-65).
// Every function is __host__ __device__ so the host verifier and the device
// kernels share one definition of the arithmetic.
//
//  * BabyBear  p = 2^31 - 2^27 + 1, Montgomery storage word = a * 2^32 mod p.
//    Products go through one 32x32->64 multiply-add and a REDC
//    (m = lo * (-p^-1) mod 2^32; (T + m p) >> 32), kept in [0, p) by an
//    unsigned-min fold.
//  * BabyBear^4 = F_p[X]/(X^4 - 11), limbs c0..c3 (each Montgomery).
//    Schoolbook with 64-bit lazy accumulation: each output limb is four
//    products summed in 64 bits (< 4 p^2), folded once by p * 2^32, then one
//    REDC. The X^4 = 11 wrap multiplies the high limbs of one operand by 11
//    with three modular doublings and two modular additions (no multiply).
//  * Goldilocks p = 2^64 - 2^32 + 1, canonical storage; 64x64->128 from 32-bit
//    partials, reduced with 2^64 = 2^32 - 1 and 2^96 = -1 (mod p).
//
// All results are fully reduced, so storage words are canonical and equal
// bit for bit to any other correct implementation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define ZS_HD __host__ __device__ __forceinline__

namespace zs {

// ---------------------------------------------------------------- BabyBear
constexpr uint32_t BB_P = 0x78000001u;
constexpr uint32_t BB_NPINV = 0x77ffffffu;  // -p^-1 mod 2^32
constexpr uint32_t BB_ONE = 0x0ffffffeu;    // 2^32 mod p (Montgomery 1)
constexpr uint32_t BB_R2 = 0x45dddde3u;     // 2^64 mod p

ZS_HD uint32_t bb_min(uint32_t a, uint32_t b) { return a < b ? a : b; }
// x in [0, 2p) -> [0, p): min(x, x - p) as unsigned.
ZS_HD uint32_t bb_fold(uint32_t x) { return bb_min(x, x - BB_P); }
ZS_HD uint32_t bb_add(uint32_t a, uint32_t b) { return bb_fold(a + b); }
ZS_HD uint32_t bb_sub(uint32_t a, uint32_t b) {
  uint32_t d = a - b;
  return bb_min(d, d + BB_P);
}
ZS_HD uint32_t bb_neg(uint32_t a) { return bb_sub(0u, a); }
ZS_HD uint32_t bb_dbl(uint32_t a) { return bb_fold(a << 1); }

// REDC of a 64-bit value T < p * 2^32 (result in [0, p)).
ZS_HD uint32_t bb_redc(uint64_t t) {
  uint32_t m = (uint32_t)t * BB_NPINV;
  uint64_t s = t + (uint64_t)m * BB_P;  // < 2 p 2^32 < 2^64
  return bb_fold((uint32_t)(s >> 32));
}
// Fold a 64-bit accumulator of up to four products (< 4 p^2 < 2 p 2^32)
// below p * 2^32 without changing its REDC residue.
ZS_HD uint64_t bb_pre(uint64_t t) {
  uint32_t hi = (uint32_t)(t >> 32);
  hi = bb_min(hi, hi - BB_P);
  return ((uint64_t)hi << 32) | (uint32_t)t;
}
ZS_HD uint32_t bb_mul(uint32_t a, uint32_t b) { return bb_redc((uint64_t)a * b); }
// 11 a: one Montgomery product with the Montgomery form of 11
// (11 * 2^32 mod p), cheaper than shift-and-add with a fold per step.
constexpr uint32_t BB_MONT11 = 0x37ffffe9u;
ZS_HD uint32_t bb_mul11(uint32_t a) { return bb_mul(a, BB_MONT11); }
ZS_HD uint32_t bb_to_mont(uint32_t v) { return bb_mul(v % BB_P, BB_R2); }
ZS_HD uint32_t bb_from_mont(uint32_t w) { return bb_redc((uint64_t)w); }
// Reduce any 32-bit word to [0, p) (same Montgomery value).
ZS_HD uint32_t bb_canon(uint32_t w) {
  w = bb_min(w, w - BB_P);
  return bb_min(w, w - BB_P);
}
ZS_HD uint32_t bb_from_int(int64_t v) {
  int64_t m = v % (int64_t)BB_P;
  if (m < 0) m += BB_P;
  return bb_to_mont((uint32_t)m);
}
ZS_HD uint32_t bb_pow(uint32_t a, uint64_t e) {
  uint32_t r = BB_ONE;
  while (e) {
    if (e & 1) r = bb_mul(r, a);
    a = bb_mul(a, a);
    e >>= 1;
  }
  return r;
}
ZS_HD uint32_t bb_inv(uint32_t a) { return bb_pow(a, BB_P - 2); }

// ------------------------------------------------------------- BabyBear^4
struct bb4 {
  uint32_t c[4];
};
ZS_HD bb4 bb4_make(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3) {
  bb4 r;
  r.c[0] = a0; r.c[1] = a1; r.c[2] = a2; r.c[3] = a3;
  return r;
}
ZS_HD bb4 bb4_zero() { return bb4_make(0, 0, 0, 0); }
ZS_HD bb4 bb4_one() { return bb4_make(BB_ONE, 0, 0, 0); }
ZS_HD bb4 bb4_embed(uint32_t a) { return bb4_make(a, 0, 0, 0); }
ZS_HD bb4 bb4_add(bb4 a, bb4 b) {
  return bb4_make(bb_add(a.c[0], b.c[0]), bb_add(a.c[1], b.c[1]), bb_add(a.c[2], b.c[2]),
                  bb_add(a.c[3], b.c[3]));
}
ZS_HD bb4 bb4_sub(bb4 a, bb4 b) {
  return bb4_make(bb_sub(a.c[0], b.c[0]), bb_sub(a.c[1], b.c[1]), bb_sub(a.c[2], b.c[2]),
                  bb_sub(a.c[3], b.c[3]));
}
ZS_HD bb4 bb4_scale(bb4 a, uint32_t s) {  // extension times base
  return bb4_make(bb_mul(a.c[0], s), bb_mul(a.c[1], s), bb_mul(a.c[2], s), bb_mul(a.c[3], s));
}
// Multiplication with the X^4 = 11 wrap folded into b's high limbs:
// b1w, b2w, b3w = 11 b1, 11 b2, 11 b3 (mod p).
ZS_HD bb4 bb4_mul_w(bb4 a, bb4 b, uint32_t b1w, uint32_t b2w, uint32_t b3w) {
  uint64_t t0 = (uint64_t)a.c[0] * b.c[0];
  t0 += (uint64_t)a.c[1] * b3w;
  t0 += (uint64_t)a.c[2] * b2w;
  t0 += (uint64_t)a.c[3] * b1w;
  uint64_t t1 = (uint64_t)a.c[0] * b.c[1];
  t1 += (uint64_t)a.c[1] * b.c[0];
  t1 += (uint64_t)a.c[2] * b3w;
  t1 += (uint64_t)a.c[3] * b2w;
  uint64_t t2 = (uint64_t)a.c[0] * b.c[2];
  t2 += (uint64_t)a.c[1] * b.c[1];
  t2 += (uint64_t)a.c[2] * b.c[0];
  t2 += (uint64_t)a.c[3] * b3w;
  uint64_t t3 = (uint64_t)a.c[0] * b.c[3];
  t3 += (uint64_t)a.c[1] * b.c[2];
  t3 += (uint64_t)a.c[2] * b.c[1];
  t3 += (uint64_t)a.c[3] * b.c[0];
  return bb4_make(bb_redc(bb_pre(t0)), bb_redc(bb_pre(t1)), bb_redc(bb_pre(t2)),
                  bb_redc(bb_pre(t3)));
}
ZS_HD bb4 bb4_mul(bb4 a, bb4 b) {
  return bb4_mul_w(a, b, bb_mul11(b.c[1]), bb_mul11(b.c[2]), bb_mul11(b.c[3]));
}
ZS_HD bool bb4_eq(bb4 a, bb4 b) {
  return a.c[0] == b.c[0] && a.c[1] == b.c[1] && a.c[2] == b.c[2] && a.c[3] == b.c[3];
}
ZS_HD bb4 bb4_pow(bb4 a, uint64_t e_lo, uint64_t e_hi) {
  bb4 r = bb4_one();
  for (int w = 1; w >= 0; --w) {
    uint64_t e = w ? e_hi : e_lo;
    for (int b = 63; b >= 0; --b) {
      r = bb4_mul(r, r);
      if ((e >> b) & 1) r = bb4_mul(r, a);
    }
  }
  return r;
}
// a^(p^4 - 2).
ZS_HD bb4 bb4_inv(bb4 a) {
  unsigned __int128 q = (unsigned __int128)BB_P * BB_P;
  unsigned __int128 e = q * q - 2;
  return bb4_pow(a, (uint64_t)e, (uint64_t)(e >> 64));
}

// -------------------------------------------------------------- Goldilocks
constexpr uint64_t GL_P = 0xFFFFFFFF00000001ull;
constexpr uint64_t GL_EPS = 0xFFFFFFFFull;  // 2^64 mod p

ZS_HD uint64_t gl_add(uint64_t a, uint64_t b) {
  uint64_t s = a + b;
  // On wrap add 2^64 mod p = eps; then one conditional subtract.
  s += (s < a) ? GL_EPS : 0;
  return s >= GL_P ? s - GL_P : s;
}
ZS_HD uint64_t gl_sub(uint64_t a, uint64_t b) {
  uint64_t d = a - b;
  return (a < b) ? d - GL_EPS : d;  // d + p mod 2^64 = d - eps
}
ZS_HD uint64_t gl_canon(uint64_t a) { return a >= GL_P ? a - GL_P : a; }
ZS_HD void gl_mul_wide(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
#ifdef __HIP_DEVICE_COMPILE__
  lo = a * b;
  hi = __umul64hi(a, b);
#else
  unsigned __int128 t = (unsigned __int128)a * b;
  lo = (uint64_t)t;
  hi = (uint64_t)(t >> 64);
#endif
}
ZS_HD uint64_t gl_reduce(uint64_t lo, uint64_t hi) {
  uint64_t hi_hi = hi >> 32, hi_lo = hi & GL_EPS;
  uint64_t t0 = lo - hi_hi;  // - 2^96 * hi_hi = + hi_hi ... subtract
  if (lo < hi_hi) t0 -= GL_EPS;
  uint64_t t1 = hi_lo * GL_EPS;  // 2^64 * hi_lo = hi_lo * (2^32 - 1)
  uint64_t t2 = t0 + t1;
  if (t2 < t0) t2 += GL_EPS;
  return t2 >= GL_P ? t2 - GL_P : t2;
}
ZS_HD uint64_t gl_mul(uint64_t a, uint64_t b) {
  uint64_t lo, hi;
  gl_mul_wide(a, b, lo, hi);
  return gl_reduce(lo, hi);
}
ZS_HD uint64_t gl_from_int(int64_t v) {
  return v >= 0 ? gl_canon((uint64_t)v) : GL_P - (uint64_t)(-v);
}
ZS_HD uint64_t gl_pow(uint64_t a, uint64_t e) {
  uint64_t r = 1;
  while (e) {
    if (e & 1) r = gl_mul(r, a);
    a = gl_mul(a, a);
    e >>= 1;
  }
  return r;
}
ZS_HD uint64_t gl_inv(uint64_t a) { return gl_pow(a, GL_P - 2); }

// -------------------------------------------------- field traits (generic code)
// Field ids shared with the C ABI.
enum FieldId { F_BABYBEAR = 1, F_BABYBEAR4 = 2, F_GOLDILOCKS = 3 };

struct BB {
  using T = uint32_t;
  using Word = uint32_t;
  static constexpr int id = F_BABYBEAR;
  static constexpr int words = 1;
  static constexpr int bytes = 4;
  static ZS_HD T zero() { return 0; }
  static ZS_HD T one() { return BB_ONE; }
  static ZS_HD T add(T a, T b) { return bb_add(a, b); }
  static ZS_HD T sub(T a, T b) { return bb_sub(a, b); }
  static ZS_HD T mul(T a, T b) { return bb_mul(a, b); }
  static ZS_HD T inv(T a) { return bb_inv(a); }
  static ZS_HD T from_int(int64_t v) { return bb_from_int(v); }
  static ZS_HD bool eq(T a, T b) { return a == b; }
  static ZS_HD T canon(T a) { return bb_canon(a); }
};

struct BB4 {
  using T = bb4;
  using Word = uint32_t;
  static constexpr int id = F_BABYBEAR4;
  static constexpr int words = 4;
  static constexpr int bytes = 16;
  static ZS_HD T zero() { return bb4_zero(); }
  static ZS_HD T one() { return bb4_one(); }
  static ZS_HD T add(T a, T b) { return bb4_add(a, b); }
  static ZS_HD T sub(T a, T b) { return bb4_sub(a, b); }
  static ZS_HD T mul(T a, T b) { return bb4_mul(a, b); }
  static ZS_HD T inv(T a) { return bb4_inv(a); }
  static ZS_HD T from_int(int64_t v) { return bb4_embed(bb_from_int(v)); }
  static ZS_HD bool eq(T a, T b) { return bb4_eq(a, b); }
  static ZS_HD T canon(T a) {
    return bb4_make(bb_canon(a.c[0]), bb_canon(a.c[1]), bb_canon(a.c[2]), bb_canon(a.c[3]));
  }
};

struct GL {
  using T = uint64_t;
  using Word = uint64_t;
  static constexpr int id = F_GOLDILOCKS;
  static constexpr int words = 1;
  static constexpr int bytes = 8;
  static ZS_HD T zero() { return 0; }
  static ZS_HD T one() { return 1; }
  static ZS_HD T add(T a, T b) { return gl_add(a, b); }
  static ZS_HD T sub(T a, T b) { return gl_sub(a, b); }
  static ZS_HD T mul(T a, T b) { return gl_mul(a, b); }
  static ZS_HD T inv(T a) { return gl_inv(a); }
  static ZS_HD T from_int(int64_t v) { return gl_from_int(v); }
  static ZS_HD bool eq(T a, T b) { return a == b; }
  static ZS_HD T canon(T a) { return gl_canon(a); }
};

// Embedding of a table field element into the challenge field.
template <class FT, class FC>
struct Lift;
template <class F>
struct Lift<F, F> {
  static ZS_HD typename F::T go(typename F::T a) { return a; }
};
template <>
struct Lift<BB, BB4> {
  static ZS_HD bb4 go(uint32_t a) { return bb4_embed(a); }
};

// Challenge-field element times table-field element (the first fold).
template <class FC, class FT>
struct MixMul {
  static ZS_HD typename FC::T go(typename FC::T r, typename FT::T x) {
    return FC::mul(r, Lift<FT, FC>::go(x));
  }
};
template <>
struct MixMul<BB4, BB> {
  static ZS_HD bb4 go(bb4 r, uint32_t x) { return bb4_scale(r, x); }
};

}  // namespace zs
