/* CPU oracle for the dense sumcheck path, in C (TEST INFRASTRUCTURE ONLY).
 *
 * The same algorithm as oracle/sumcheck.py (the Python restatement), sized for
 * large n: tests use it as the parity checker at sizes the Python oracle cannot
 * reach, and bench.py times it as the CPU baseline. Nothing in the product links
 * or loads it.
 *
 * Restated reference behaviour (file:line into the reference):
 *   - StandardRound(ProductSummand(d)) rounds, MSB split/fold, natural domain:
 *     zorch/sumcheck/prover.py:157-174, zorch/sumcheck/domain.py:131-137,
 *     158-163,176-205,226-233.
 *   - EqPolyRound with factored suffix eq tables: zorch/sumcheck/eq/eq_poly.py:
 *     63-111,144-221; zorch/poly/eq.py:16-26,55-62,130-214.
 *   - Claim reduction by Lagrange interpolation over {0..d}:
 *     zorch/sumcheck/reduce.py:27-52, zorch/poly/univariate.py:36-65.
 *   - SHA-256 field transcript hop observe(msg) + sample(L):
 *     zorch/sha256_field_transcript.py:100-114,189-196,251-259;
 *     zorch/challenge.py:80-106; zorch/byte_transcript.py:37-54.
 *
 * Arithmetic is on canonical residues held in uint64 limbs; the Montgomery
 * storage codec of BabyBear (word = a * 2^32 mod p) is applied only at the
 * boundary. Field ids: 1 = BabyBear, 2 = BabyBear^4 (X^4 - 11), 3 = Goldilocks.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define BB_P 2013265921ull
#define GL_P 0xFFFFFFFF00000001ull

typedef struct { uint64_t c[4]; } fe;

typedef struct {
  int id;        /* 1, 2, 3 */
  int limbs;     /* 1 or 4 */
  uint64_t p;
} field_t;

static field_t mkfield(int id) {
  field_t f;
  f.id = id;
  f.limbs = (id == 2) ? 4 : 1;
  f.p = (id == 3) ? GL_P : BB_P;
  return f;
}

static inline uint64_t gl_reduce128(unsigned __int128 x) {
  /* 2^64 = 2^32 - 1 (mod p), 2^96 = -1 (mod p). */
  uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
  uint64_t hi_hi = hi >> 32, hi_lo = hi & 0xFFFFFFFFull;
  uint64_t t0 = lo - hi_hi;
  if (lo < hi_hi) t0 -= 0xFFFFFFFFull; /* borrow: add p = subtract (2^32-1) mod 2^64 */
  uint64_t t1 = hi_lo * 0xFFFFFFFFull;
  uint64_t t2 = t0 + t1;
  if (t2 < t0) t2 += 0xFFFFFFFFull;  /* carry */
  if (t2 >= GL_P) t2 -= GL_P;
  return t2;
}

static inline uint64_t pmul(const field_t* f, uint64_t a, uint64_t b) {
  if (f->id == 3) return gl_reduce128((unsigned __int128)a * b);
  return (a * b) % BB_P;
}
static inline uint64_t padd(const field_t* f, uint64_t a, uint64_t b) {
  uint64_t s = a + b;
  if (f->id == 3) { if (s < a || s >= GL_P) s -= GL_P; return s; }
  return s >= BB_P ? s - BB_P : s;
}
static inline uint64_t psub(const field_t* f, uint64_t a, uint64_t b) {
  return a >= b ? a - b : a + (f->p - b);
}

static inline fe fadd(const field_t* f, fe a, fe b) {
  fe r = {{0, 0, 0, 0}};
  for (int i = 0; i < f->limbs; ++i) r.c[i] = padd(f, a.c[i], b.c[i]);
  return r;
}
static inline fe fsub(const field_t* f, fe a, fe b) {
  fe r = {{0, 0, 0, 0}};
  for (int i = 0; i < f->limbs; ++i) r.c[i] = psub(f, a.c[i], b.c[i]);
  return r;
}
static inline fe fmul(const field_t* f, fe a, fe b) {
  fe r = {{0, 0, 0, 0}};
  if (f->limbs == 1) { r.c[0] = pmul(f, a.c[0], b.c[0]); return r; }
  uint64_t acc[7] = {0};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) acc[i + j] = (acc[i + j] + a.c[i] * b.c[j]) % BB_P;
  for (int k = 0; k < 4; ++k) r.c[k] = (acc[k] + (k + 4 < 7 ? 11 * acc[k + 4] : 0)) % BB_P;
  return r;
}
static inline fe fint(const field_t* f, int64_t v) {
  fe r = {{0, 0, 0, 0}};
  if (f->id == 3) { r.c[0] = v >= 0 ? (uint64_t)v : GL_P - (uint64_t)(-v); return r; }
  int64_t m = v % (int64_t)BB_P;
  if (m < 0) m += (int64_t)BB_P;
  r.c[0] = (uint64_t)m;
  return r;
}
static fe fpow(const field_t* f, fe a, const uint64_t* e, int nwords) {
  fe r = fint(f, 1);
  for (int w = nwords - 1; w >= 0; --w)
    for (int b = 63; b >= 0; --b) {
      r = fmul(f, r, r);
      if ((e[w] >> b) & 1) r = fmul(f, r, a);
    }
  return r;
}
static fe finv(const field_t* f, fe a) {
  if (f->limbs == 1) { uint64_t e = f->p - 2; return fpow(f, a, &e, 1); }
  /* p^4 - 2 as 128-bit little-endian words. */
  unsigned __int128 q = (unsigned __int128)BB_P * BB_P;
  unsigned __int128 q4 = q * q - 2;
  uint64_t e[2] = {(uint64_t)q4, (uint64_t)(q4 >> 64)};
  return fpow(f, a, e, 2);
}
static inline int feq(const field_t* f, fe a, fe b) {
  for (int i = 0; i < f->limbs; ++i) if (a.c[i] != b.c[i]) return 0;
  return 1;
}

/* --- storage codec ---------------------------------------------------- */
static const uint64_t R_MOD_BB = (1ull << 32) % BB_P;
static uint64_t RINV_BB = 0;
static void init_rinv(void) {
  if (RINV_BB) return;
  field_t b = mkfield(1);
  fe r = {{R_MOD_BB, 0, 0, 0}};
  RINV_BB = finv(&b, r).c[0];
}
/* Storage word(s) at `src` (element index i) -> value. */
static inline fe load_el(const field_t* f, const void* src, size_t i) {
  fe r = {{0, 0, 0, 0}};
  if (f->id == 3) { r.c[0] = ((const uint64_t*)src)[i] % GL_P; return r; }
  const uint32_t* w = (const uint32_t*)src + i * f->limbs;
  for (int k = 0; k < f->limbs; ++k) r.c[k] = ((uint64_t)w[k] % BB_P) * RINV_BB % BB_P;
  return r;
}
static inline void store_el(const field_t* f, void* dst, size_t i, fe v) {
  if (f->id == 3) { ((uint64_t*)dst)[i] = v.c[0]; return; }
  uint32_t* w = (uint32_t*)dst + i * f->limbs;
  for (int k = 0; k < f->limbs; ++k) w[k] = (uint32_t)((v.c[k] << 32) % BB_P);
}

/* --- SHA-256 streaming ------------------------------------------------- */
typedef struct {
  uint32_t h[8];
  uint8_t pending[64];
  uint64_t pending_len;
  uint64_t total_len;
} oref_sha_state;

static const uint32_t K[64] = {
  0x428a2f98,0x71374491,0xb5c0fbcf,0xe9b5dba5,0x3956c25b,0x59f111f1,0x923f82a4,0xab1c5ed5,
  0xd807aa98,0x12835b01,0x243185be,0x550c7dc3,0x72be5d74,0x80deb1fe,0x9bdc06a7,0xc19bf174,
  0xe49b69c1,0xefbe4786,0x0fc19dc6,0x240ca1cc,0x2de92c6f,0x4a7484aa,0x5cb0a9dc,0x76f988da,
  0x983e5152,0xa831c66d,0xb00327c8,0xbf597fc7,0xc6e00bf3,0xd5a79147,0x06ca6351,0x14292967,
  0x27b70a85,0x2e1b2138,0x4d2c6dfc,0x53380d13,0x650a7354,0x766a0abb,0x81c2c92e,0x92722c85,
  0xa2bfe8a1,0xa81a664b,0xc24b8b70,0xc76c51a3,0xd192e819,0xd6990624,0xf40e3585,0x106aa070,
  0x19a4c116,0x1e376c08,0x2748774c,0x34b0bcb5,0x391c0cb3,0x4ed8aa4a,0x5b9cca4f,0x682e6ff3,
  0x748f82ee,0x78a5636f,0x84c87814,0x8cc70208,0x90befffa,0xa4506ceb,0xbef9a3f7,0xc67178f2};

#define ROR(x, n) (((x) >> (n)) | ((x) << (32 - (n))))
static void sha_compress(uint32_t h[8], const uint8_t blk[64]) {
  uint32_t w[64];
  for (int t = 0; t < 16; ++t)
    w[t] = ((uint32_t)blk[4 * t] << 24) | ((uint32_t)blk[4 * t + 1] << 16) |
           ((uint32_t)blk[4 * t + 2] << 8) | blk[4 * t + 3];
  for (int t = 16; t < 64; ++t) {
    uint32_t s0 = ROR(w[t - 15], 7) ^ ROR(w[t - 15], 18) ^ (w[t - 15] >> 3);
    uint32_t s1 = ROR(w[t - 2], 17) ^ ROR(w[t - 2], 19) ^ (w[t - 2] >> 10);
    w[t] = w[t - 16] + s0 + w[t - 7] + s1;
  }
  uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
  for (int t = 0; t < 64; ++t) {
    uint32_t t1 = hh + (ROR(e, 6) ^ ROR(e, 11) ^ ROR(e, 25)) + ((e & f) ^ (~e & g)) + K[t] + w[t];
    uint32_t t2 = (ROR(a, 2) ^ ROR(a, 13) ^ ROR(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
    hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
  }
  h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
}
static void sha_absorb(oref_sha_state* s, const uint8_t* data, size_t len) {
  for (size_t i = 0; i < len; ++i) {
    s->pending[s->pending_len++] = data[i];
    if (s->pending_len == 64) { sha_compress(s->h, s->pending); s->pending_len = 0; }
  }
  s->total_len += len;
}
static void sha_digest(const oref_sha_state* s0, uint8_t out[32]) {
  oref_sha_state s = *s0;
  uint64_t bits = s.total_len * 8;
  uint8_t pad = 0x80, z = 0;
  sha_absorb(&s, &pad, 1);
  while (s.pending_len != 56) sha_absorb(&s, &z, 1);
  uint8_t lb[8];
  for (int i = 0; i < 8; ++i) lb[i] = (uint8_t)(bits >> (56 - 8 * i));
  sha_absorb(&s, lb, 8);
  for (int i = 0; i < 8; ++i) {
    out[4 * i] = (uint8_t)(s.h[i] >> 24); out[4 * i + 1] = (uint8_t)(s.h[i] >> 16);
    out[4 * i + 2] = (uint8_t)(s.h[i] >> 8); out[4 * i + 3] = (uint8_t)s.h[i];
  }
}
static void le64(uint8_t* o, uint64_t v) { for (int i = 0; i < 8; ++i) o[i] = (uint8_t)(v >> (8 * i)); }

/* observe(msg bytes, count) then sample(L words of item bytes). */
static void fs_hop(oref_sha_state* s, const uint8_t* msg, size_t msg_len, size_t item,
                   size_t limbs, uint8_t* out) {
  uint8_t fr[10];
  fr[0] = 3; fr[1] = 2; le64(fr + 2, msg_len / item);
  sha_absorb(s, fr, 10);
  sha_absorb(s, msg, msg_len);
  fr[0] = 4; fr[1] = 2; le64(fr + 2, limbs);
  sha_absorb(s, fr, 10);
  size_t nbytes = limbs * item, got = 0;
  for (uint64_t ctr = 0; got < nbytes; ++ctr) {
    oref_sha_state t = *s;
    uint8_t c8[8], dg[32];
    le64(c8, ctr);
    sha_absorb(&t, c8, 8);
    sha_digest(&t, dg);
    size_t take = nbytes - got < 32 ? nbytes - got : 32;
    memcpy(out + got, dg, take);
    got += take;
  }
  sha_absorb(s, out, nbytes);
}

/* --- Lagrange over {0..d} ------------------------------------------------ */
static fe lagrange(const field_t* f, const fe* ev, int n, fe r) {
  fe out = fint(f, 0);
  for (int k = 0; k < n; ++k) {
    fe num = fint(f, 1), den = fint(f, 1);
    for (int j = 0; j < n; ++j) {
      if (j == k) continue;
      num = fmul(f, num, fsub(f, r, fint(f, j)));
      den = fmul(f, den, fint(f, k - j));
    }
    out = fadd(f, out, fmul(f, ev[k], fmul(f, num, finv(f, den))));
  }
  return out;
}

/* Suffix eq table eq(w[lo..hi), .) with w[lo] on the MSB; out has 2^(hi-lo). */
static void eq_suffix(const field_t* f, const fe* w, int lo, int hi, fe* out) {
  out[0] = fint(f, 1);
  size_t len = 1;
  for (int j = hi - 1; j >= lo; --j) {
    for (size_t x = 0; x < len; ++x) {
      fe hh = fmul(f, out[x], w[j]);
      out[x + len] = hh;
      out[x] = fsub(f, out[x], hh);
    }
    len *= 2;
  }
}

#define MAXK 8
#define MAXD 10

/* Prove. Field ids as above. `tab_field` is the storage field of `tables`
 * (k tables of 2^n elements, table-major); `chal_field` is the challenge /
 * message field; `tr_field` the transcript word field. `eq`=1 selects
 * EqPolyProver over the summand ProductSummand(degree) (messages of degree+2
 * values), else StandardRound(ProductSummand(degree)) (degree+1 values).
 * Returns 0 on success, -1 on bad arguments. */
int oref_prove(int tab_field, int chal_field, int tr_field, int eq, int degree, int k, int n,
               const void* tables, const void* eq_point, const void* claim_in,
               oref_sha_state* tr, void* msgs_out, void* point_out, void* value_out,
               void* final_out, int threads) {
  if (k < 1 || k > MAXK || degree < 1 || degree + 2 > MAXD || n < 1 || n > 34) return -1;
  init_rinv();
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  field_t FT = mkfield(tab_field), F = mkfield(chal_field), T = mkfield(tr_field);
  size_t N = (size_t)1 << n;
  size_t item = (tr_field == 3) ? 8 : (tr_field == 2 ? 16 : 4);
  size_t chal_bytes = (chal_field == 3) ? 8 : (chal_field == 2 ? 16 : 4);
  size_t limbs = chal_bytes / item;
  (void)T;
  fe* st = (fe*)malloc(sizeof(fe) * N * (size_t)k);
  if (!st) return -2;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N * (size_t)k; ++i) st[i] = load_el(&FT, tables, i);
  fe claim = load_el(&F, claim_in, 0);
  int npts = eq ? degree + 2 : degree + 1;
  fe *w = NULL, *eql = NULL, *eqr_full = NULL;
  int l_half = n / 2;
  fe eq_prev = fint(&F, 1);
  if (eq) {
    w = (fe*)malloc(sizeof(fe) * n);
    for (int i = 0; i < n; ++i) w[i] = load_el(&F, eq_point, i);
    eql = (fe*)malloc(sizeof(fe) * ((size_t)1 << l_half));
    eqr_full = (fe*)malloc(sizeof(fe) * ((size_t)1 << (n - l_half)));
  }
  uint8_t msgbuf[MAXD * 16], sq[64];
  size_t len = N;
  for (int rnd = 0; rnd < n; ++rnd) {
    size_t h = len / 2;
    fe acc[MAXD];
    for (int t = 0; t < npts; ++t) acc[t] = fint(&F, 0);
    /* eq weights for this round (eq_poly.py:144-187); round index i = rnd+1. */
    int i1 = rnd + 1, rest = n - i1;
    int use_l = 0;
    size_t nxr = 1;
    if (eq) {
      if (rest == 0) { eqr_full[0] = fint(&F, 1); nxr = 1; }
      else if (i1 < l_half) {
        eq_suffix(&F, w, i1, l_half, eql);
        eq_suffix(&F, w, l_half, n, eqr_full);
        use_l = 1; nxr = (size_t)1 << (n - l_half);
      } else {
        eq_suffix(&F, w, i1, n, eqr_full);
        nxr = (size_t)1 << rest;
      }
    }
#pragma omp parallel
    {
      fe loc[MAXD];
      for (int t = 0; t < npts; ++t) loc[t] = fint(&F, 0);
#pragma omp for schedule(static)
      for (size_t j = 0; j < h; ++j) {
        fe v[MAXK], d[MAXK];
        for (int q = 0; q < k; ++q) {
          fe lo = st[(size_t)q * N + j], hi = st[(size_t)q * N + j + h];
          v[q] = lo; d[q] = fsub(&F, hi, lo);
        }
        fe wt = fint(&F, 1);
        if (eq) wt = use_l ? fmul(&F, eql[j / nxr], eqr_full[j % nxr]) : eqr_full[j];
        for (int t = 0; t < npts; ++t) {
          fe pr = v[0];
          for (int q = 1; q < k; ++q) pr = fmul(&F, pr, v[q]);
          if (eq) pr = fmul(&F, pr, wt);
          loc[t] = fadd(&F, loc[t], pr);
          for (int q = 0; q < k; ++q) v[q] = fadd(&F, v[q], d[q]);
        }
      }
#pragma omp critical
      for (int t = 0; t < npts; ++t) acc[t] = fadd(&F, acc[t], loc[t]);
    }
    fe msg[MAXD];
    fe wi = fint(&F, 0);
    if (eq) {
      wi = w[rnd];
      fe l1 = fmul(&F, eq_prev, wi);
      fe l0 = fsub(&F, eq_prev, l1);
      fe ld = fsub(&F, l1, l0);
      fe lt = l0;
      for (int t = 0; t < npts; ++t) { msg[t] = fmul(&F, lt, acc[t]); lt = fadd(&F, lt, ld); }
    } else {
      for (int t = 0; t < npts; ++t) msg[t] = acc[t];
    }
    for (int t = 0; t < npts; ++t) {
      store_el(&F, msgs_out, (size_t)rnd * npts + t, msg[t]);
      store_el(&F, msgbuf, t, msg[t]);
    }
    fs_hop(tr, msgbuf, (size_t)npts * chal_bytes, item, limbs, sq);
    memcpy((uint8_t*)point_out + (size_t)rnd * chal_bytes, sq, chal_bytes);
    fe r = load_el(&F, sq, 0);
    claim = lagrange(&F, msg, npts, r);
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < h; ++j)
      for (int q = 0; q < k; ++q) {
        fe lo = st[(size_t)q * N + j], hi = st[(size_t)q * N + j + h];
        st[(size_t)q * N + j] = fadd(&F, lo, fmul(&F, r, fsub(&F, hi, lo)));
      }
    if (eq) {
      fe one = fint(&F, 1);
      fe ef = fadd(&F, fmul(&F, r, wi), fmul(&F, fsub(&F, one, r), fsub(&F, one, wi)));
      eq_prev = fmul(&F, eq_prev, ef);
    }
    len = h;
  }
  store_el(&F, value_out, 0, claim);
  if (final_out) {
    for (int q = 0; q < k; ++q) store_el(&F, final_out, q, st[(size_t)q * N]);
    if (eq) store_el(&F, final_out, k, eq_prev);
  }
  free(st); free(w); free(eql); free(eqr_full);
  return 0;
}

/* Claim helper: sum_x eq(w, x)^[eq] * prod_k T_k[x], in the challenge field. */
int oref_claim(int tab_field, int chal_field, int eq, int k, int n, const void* tables,
               const void* eq_point, void* out, int threads) {
  init_rinv();
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  field_t FT = mkfield(tab_field), F = mkfield(chal_field);
  size_t N = (size_t)1 << n;
  fe* eqt = NULL;
  if (eq) {
    fe* w = (fe*)malloc(sizeof(fe) * n);
    for (int i = 0; i < n; ++i) w[i] = load_el(&F, eq_point, i);
    eqt = (fe*)malloc(sizeof(fe) * N);
    eq_suffix(&F, w, 0, n, eqt);
    free(w);
  }
  fe acc = fint(&F, 0);
#pragma omp parallel
  {
    fe loc = fint(&F, 0);
#pragma omp for schedule(static)
    for (size_t x = 0; x < N; ++x) {
      fe pr = load_el(&FT, tables, x);
      for (int q = 1; q < k; ++q) pr = fmul(&F, pr, load_el(&FT, tables, (size_t)q * N + x));
      if (eq) pr = fmul(&F, pr, eqt[x]);
      loc = fadd(&F, loc, pr);
    }
#pragma omp critical
    acc = fadd(&F, acc, loc);
  }
  store_el(&F, out, 0, acc);
  free(eqt);
  return 0;
}

/* Transcript helpers for cross-checks. */
void oref_sha_init(oref_sha_state* s) {
  static const uint32_t H0[8] = {0x6a09e667,0xbb67ae85,0x3c6ef372,0xa54ff53a,
                                 0x510e527f,0x9b05688c,0x1f83d9ab,0x5be0cd19};
  memcpy(s->h, H0, sizeof H0);
  memset(s->pending, 0, 64);
  s->pending_len = 0;
  s->total_len = 0;
}
void oref_sha_absorb(oref_sha_state* s, const uint8_t* d, size_t n) { sha_absorb(s, d, n); }
void oref_sha_digest(const oref_sha_state* s, uint8_t out[32]) { sha_digest(s, out); }
