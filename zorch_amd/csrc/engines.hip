// Univariate-skip and small-value engines: their first-phase kernels and the
// host drivers that hand the rest of the stage to the dense prover
// (zs_prove_univariate_skip, zs_prove_small_value in include/zorch_amd.h).
#include <string.h>

#include <memory>
#include <vector>

#include "engines.h"
#include "instance.h"
#include "runtime.h"

namespace zs {

// ------------------------------------------------------------- shared pieces
// V_q[x] = sum_b w[b] P_q[b L + x] for x < L (q < k): the prefix of `blog`
// high variables of each table bound at the weights w (2^blog challenge-field
// elements). The univariate skip binds its prism with the Lagrange basis of
// the subgroup at r0 (univariate_skip.py:124-133); the small-value transition
// refolds the factors with eq(r[<l0], .) (small_value.py:180-181,
// sqrt_space.py:40-48).
template <class FT, class FC>
__global__ __launch_bounds__(256) void weighted_refold_kernel(const typename FT::T* in,
                                                              size_t in_stride, uint32_t k,
                                                              uint32_t blog, uint32_t llog,
                                                              const typename FC::T* w,
                                                              typename FC::T* out,
                                                              size_t out_stride) {
  const size_t L = (size_t)1 << llog;
  const size_t total = (size_t)k << llog;
  const uint32_t B = 1u << blog;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total;
       i += (size_t)gridDim.x * 256) {
    const size_t q = i >> llog, x = i & (L - 1);
    const typename FT::T* col = in + q * in_stride + x;
    WeightedSum<FC, FT> s;
    s.init();
    for (uint32_t b = 0; b < B; ++b) s.add(w[b], col[(size_t)b << llog]);
    out[q * out_stride + x] = s.get();
  }
}

// Host: slot-row totals (NSLOT rows of `row` words) of element i (W words).
template <class F>
typename F::T slot_total(const uint64_t* rows, size_t row, size_t i);
template <>
uint32_t slot_total<BB>(const uint64_t* rows, size_t row, size_t i) {
  unsigned __int128 t = 0;
  for (int s = 0; s < NSLOT; ++s) t += rows[(size_t)s * row + i];
  return (uint32_t)(t % BB_P);  // sums of Montgomery words are Montgomery words of the sum
}
template <>
uint64_t slot_total<GL>(const uint64_t* rows, size_t row, size_t i) {
  unsigned __int128 lo = 0, hi = 0;
  for (int s = 0; s < NSLOT; ++s) {
    lo += rows[(size_t)s * row + 2 * i];
    hi += rows[(size_t)s * row + 2 * i + 1];
  }
  return GL::add((uint64_t)(lo % GL_P), GL::mul((uint64_t)(hi % GL_P), (uint64_t)1 << 32));
}

template <class F>
typename F::T fpow(typename F::T a, uint64_t e) {
  typename F::T r = F::one();
  while (e) {
    if (e & 1) r = F::mul(r, a);
    a = F::mul(a, a);
    e >>= 1;
  }
  return r;
}

// A primitive 2^k-th root of unity: the multiplicative generator (BabyBear
// 31, Goldilocks 7) raised to (p - 1) / 2^k. The reference takes its
// subgroup from FRX's NTT (domain.py:245-266), whose root is not visible
// here; every identity the round relies on holds for any primitive root.
template <class F>
typename F::T subgroup_root(uint32_t k);
template <>
uint32_t subgroup_root<BB>(uint32_t k) {
  return fpow<BB>(BB::from_int(31), (uint64_t)(BB_P - 1) >> k);
}
template <>
uint64_t subgroup_root<GL>(uint32_t k) {
  return fpow<GL>(GL::from_int(7), (GL_P - 1) >> k);
}

// Device buffers of one engine call from the context's pool, released at
// scope exit (after the caller's stream has synchronised).
struct PoolBlock {
  zs_ctx* ctx = nullptr;
  void* p = nullptr;
  ~PoolBlock() {
    if (p) ctx->pool->release(p);
  }
  int alloc(zs_ctx* c, size_t bytes) {
    ctx = c;
    return c->pool->alloc(bytes, &p);
  }
};

// The dense rounds after an engine's first phase: a whole stage on
// challenge-field tables already in HBM, continuing the transcript.
static int prove_tail(zs_ctx* ctx, const zs_prove_args& base, uint32_t chal_field, uint32_t mode,
                      const void* tables, uint32_t rounds, const void* claim,
                      const void* eq_point, const void* eq_mass, zs_transcript* t,
                      zs_proof* out) {
  zs_prove_args a = base;
  a.table_field = chal_field;
  a.chal_field = chal_field;
  a.mode = mode;
  a.tables = tables;
  a.tables_on_device = 1;
  a.rounds = rounds;
  a.claim = claim;
  a.eq_point = eq_point;
  a.eq_mass = eq_mass;
  a.refold_rounds = 0;
  uint32_t npts = 0;
  int rc = validate_args(a, t, &npts);
  if (rc) return rc;
  std::unique_ptr<Instance> inst(make_instance(chal_field, chal_field, a.num_tables, npts,
                                               mode == ZS_MODE_EQ));
  if (!inst) return fail(ZS_EINVAL, "unsupported field / degree combination");
  if ((rc = inst->setup(ctx, a, *t))) return rc;
  if ((rc = inst->enqueue(0, rounds, ctx->stream))) return rc;
  if ((rc = inst->download(ctx->stream))) return rc;
  ZS_HIP(hipStreamSynchronize(ctx->stream));
  return inst->collect(out, t);
}

// ---------------------------------------------------- univariate skip round 0
// Round 0 of the skip (univariate_skip.py:65-101): the skip high variables of
// each factor index the order-D subgroup, so column x (the 2^hn low bits)
// holds D values of a degree < D univariate. The product of the K factors
// has degree K (D - 1) < M; its values on the order-M subgroup are products
// of the factors' values there, each a fixed linear map of the column
// (iNTT over D, zero-pad, NTT over M: the M x D matrix `wmat`, built on the
// host). Thread (column c, node m) of a block forms one such product per
// column it visits and sums them; the block adds its M sums to a slot row.
// The host interpolates the M sums back to the message's coefficients. All
// base-field arithmetic.
constexpr uint32_t SKIP_MAX_W = 4096;   // M x D matrix entries
constexpr uint32_t SKIP_TILE = 2048;    // K x D x (256 / M) column values of a step
constexpr uint32_t SKIP_ROW = 512;      // 64-bit words of a slot row (M <= 256, 2 words)

template <class FT, int K>
__global__ __launch_bounds__(256) void skip_round0_kernel(const typename FT::T* in,
                                                          size_t in_stride, uint32_t hn_log,
                                                          uint32_t dlog, uint32_t mlog,
                                                          const typename FT::T* wmat,
                                                          uint64_t* slots) {
  using T = typename FT::T;
  __shared__ T w[SKIP_MAX_W];
  __shared__ T tile[SKIP_TILE];  // [K][D][cols] column values of the current step
  __shared__ T red[256];
  const uint32_t tid = threadIdx.x;
  const uint32_t D = 1u << dlog, M = 1u << mlog;
  for (uint32_t i = tid; i < (M << dlog); i += 256) w[i] = wmat[i];
  const uint32_t m = tid & (M - 1), c = tid >> mlog;
  const uint32_t clog = 8u - mlog, cols = 1u << clog;
  const size_t hn = (size_t)1 << hn_log;
  const T* wm = w + ((size_t)m << dlog);
  const uint32_t nt = (uint32_t)K << (dlog + clog);  // tile entries
  T acc = FT::zero();
  // Every thread runs every step (the tile load is shared): the steps are
  // the block's column groups, grid-strided.
  for (size_t x0 = (size_t)blockIdx.x * cols; x0 < hn; x0 += (size_t)gridDim.x * cols) {
    __syncthreads();
    for (uint32_t i = tid; i < nt; i += 256) {
      const uint32_t cc = i & (cols - 1), qb = i >> clog;  // qb = q D + b
      const uint32_t q = qb >> dlog, b = qb & (D - 1);
      const size_t x = x0 + cc;
      tile[i] = x < hn ? in[(size_t)q * in_stride + ((size_t)b << hn_log) + x] : FT::zero();
    }
    __syncthreads();
    if (x0 + c < hn) {
      T prod = FT::one();
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const T* col = tile + ((size_t)q << (dlog + clog)) + c;
        T e = FT::zero();
        for (uint32_t b = 0; b < D; ++b) e = FT::add(e, FT::mul(wm[b], col[b << clog]));
        prod = q == 0 ? e : FT::mul(prod, e);
      }
      acc = FT::add(acc, prod);
    }
  }
  red[tid] = acc;
  __syncthreads();
  if (tid < M) {
    T s = FT::zero();
    for (uint32_t i = tid; i < 256; i += M) s = FT::add(s, red[i]);
    constexpr int W = SlotOps<FT>::W;
    uint64_t* row = slots + (size_t)(blockIdx.x & (NSLOT - 1)) * SKIP_ROW;
#pragma unroll
    for (int c2 = 0; c2 < W; ++c2)
      __hip_atomic_fetch_add(row + tid * W + c2, (uint64_t)SlotOps<FT>::word(s, c2),
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Transform over the order-n subgroup with root `w` (natural order):
// out[j] = sum_i v[i] w^(ij).
template <class F>
std::vector<typename F::T> host_dft(const std::vector<typename F::T>& v, typename F::T w) {
  const size_t n = v.size();
  std::vector<typename F::T> out(n, F::zero());
  typename F::T wj = F::one();
  for (size_t j = 0; j < n; ++j) {
    typename F::T x = F::one(), acc = F::zero();
    for (size_t i = 0; i < n; ++i) {
      acc = F::add(acc, F::mul(v[i], x));
      x = F::mul(x, wj);
    }
    out[j] = acc;
    wj = F::mul(wj, w);
  }
  return out;
}

// Round 0 of the skip and the binding of the prism at r0: the message, the
// raw challenge and s0(r0) to the host; the bound tables (K x 2^(n - skip)
// challenge-field elements, `v_dev`, or a buffer of the call when null) on
// the device. Ends with the stream synchronised.
template <class FT, class FC, int K>
int skip_bind_typed(zs_ctx* ctx, const zs_prove_args& a, uint32_t skip, zs_transcript* t,
                    void* msg0_out, uint8_t* raw_out, typename FC::T* reduced_out,
                    typename FC::T* v_dev, PoolBlock& blk) {
  using TT = typename FT::T;
  using TC = typename FC::T;
  const uint32_t n = a.rounds, hn_log = n - skip;
  const uint32_t D = 1u << skip;
  const uint32_t d0 = K * (D - 1);
  uint32_t mlog = 0;
  while ((1u << mlog) < d0 + 1) ++mlog;
  const uint32_t M = 1u << mlog;
  if (M > 256 || (size_t)M * D > SKIP_MAX_W || (size_t)K * D * (256 / M) > SKIP_TILE)
    return fail(ZS_EINVAL, "skip_rounds too large: the round-0 message needs " +
                               std::to_string(M) + " subgroup values (at most 256, and " +
                               "at most 4096 extension-matrix entries)");
  const size_t N = (size_t)1 << n;
  const size_t hn = (size_t)1 << hn_log;
  hipStream_t s = ctx->stream;
  // Device block: staged input | extension matrix | slot rows | weights | V.
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off = align256(off + bytes); return o; };
  const size_t o_in = a.tables_on_device ? 0 : take((size_t)K * N * sizeof(TT));
  const size_t o_w = take((size_t)M * D * sizeof(TT));
  const size_t o_slots = take((size_t)NSLOT * SKIP_ROW * sizeof(uint64_t));
  const size_t o_lam = take((size_t)D * sizeof(TC));
  const size_t o_v = v_dev ? 0 : take((size_t)K * hn * sizeof(TC));
  int rc = blk.alloc(ctx, off);
  if (rc) return rc;
  uint8_t* b = (uint8_t*)blk.p;
  const TT* in = a.tables_on_device ? (const TT*)a.tables : (const TT*)(b + o_in);
  if (!a.tables_on_device)
    ZS_HIP(hipMemcpyAsync(b + o_in, a.tables, (size_t)K * N * sizeof(TT), hipMemcpyHostToDevice, s));
  // W[m][b] = sum_{j<D} wM^(m j) (1/D) wD^(-b j): column values -> values on
  // the order-M subgroup (zero-padded coefficients).
  const TT wD = subgroup_root<FT>(skip), wM = subgroup_root<FT>(mlog);
  const TT wD_inv = FT::inv(wD), d_inv = FT::inv(FT::from_int(D));
  std::vector<TT> W((size_t)M * D);
  for (uint32_t m = 0; m < M; ++m)
    for (uint32_t bb = 0; bb < D; ++bb) {
      TT acc = FT::zero();
      const TT step = FT::mul(fpow<FT>(wM, m), fpow<FT>(wD_inv, bb));
      TT x = FT::one();
      for (uint32_t j = 0; j < D; ++j) {
        acc = FT::add(acc, x);
        x = FT::mul(x, step);
      }
      W[(size_t)m * D + bb] = FT::mul(acc, d_inv);
    }
  ZS_HIP(hipMemcpyAsync(b + o_w, W.data(), W.size() * sizeof(TT), hipMemcpyHostToDevice, s));
  ZS_HIP(hipMemsetAsync(b + o_slots, 0, (size_t)NSLOT * SKIP_ROW * sizeof(uint64_t), s));
  size_t grid = (hn * M + 255) / 256;
  if (grid > 2048) grid = 2048;
  if (grid < 1) grid = 1;
  hipEvent_t ev = nullptr;
  if (ctx->profile) {
    ev = ctx->event();
    (void)hipEventRecord(ev, s);
  }
  skip_round0_kernel<FT, K><<<(uint32_t)grid, 256, 0, s>>>(in, N, hn_log, skip, mlog,
                                                           (const TT*)(b + o_w),
                                                           (uint64_t*)(b + o_slots));
  ZS_HIP(hipGetLastError());
  if (ev) {
    hipEvent_t stop = ctx->event();
    (void)hipEventRecord(stop, s);
    ctx->records.push_back(KernelRecord{0, KREC_ENGINE, (uint32_t)grid,
                                        (uint64_t)K * N * sizeof(TT), ev, stop});
  }
  std::vector<uint64_t> rows((size_t)NSLOT * SKIP_ROW);
  ZS_HIP(hipMemcpyAsync(rows.data(), b + o_slots, rows.size() * sizeof(uint64_t),
                        hipMemcpyDeviceToHost, s));
  ZS_HIP(hipStreamSynchronize(s));
  std::vector<TT> vals(M);
  for (uint32_t m = 0; m < M; ++m) vals[m] = slot_total<FT>(rows.data(), SKIP_ROW, m);
  // Coefficients: iNTT over the order-M subgroup, the first d0 + 1 kept.
  std::vector<TT> coef = host_dft<FT>(vals, FT::inv(wM));
  const TT m_inv = FT::inv(FT::from_int(M));
  std::vector<TT> msg0(d0 + 1);
  for (uint32_t j = 0; j <= d0; ++j) msg0[j] = FT::mul(coef[j], m_inv);
  // Fiat-Shamir: observe the base-field message, sample the challenge.
  const size_t item = t->item_bytes;
  std::vector<uint8_t> raw(sizeof(TC));
  if (!transcript_hop(t, reinterpret_cast<const uint8_t*>(msg0.data()), msg0.size() * sizeof(TT),
                      item, sizeof(TC) / item, raw.data()))
    return fail(ZS_EINVAL, "transcript stream exceeds the BLAKE3 subtree stack");
  TC r0;
  memcpy(&r0, raw.data(), sizeof(TC));
  r0 = FC::canon(r0);
  // s0(r0) and the Lagrange basis of the subgroup at r0:
  // lambda_b = (1/D) sum_j wD^(-b j) r0^j.
  TC reduced = FC::zero(), rp = FC::one();
  std::vector<TC> rpow(D);
  for (uint32_t j = 0; j <= d0; ++j) {
    reduced = FC::add(reduced, FC::mul(Lift<FT, FC>::go(msg0[j]), rp));
    if (j < D) rpow[j] = rp;
    rp = FC::mul(rp, r0);
  }
  std::vector<TC> lam(D);
  for (uint32_t bb = 0; bb < D; ++bb) {
    TC acc = FC::zero();
    const TT step = fpow<FT>(wD_inv, bb);
    TT x = FT::one();
    for (uint32_t j = 0; j < D; ++j) {
      acc = FC::add(acc, FC::mul(rpow[j], Lift<FT, FC>::go(x)));
      x = FT::mul(x, step);
    }
    lam[bb] = FC::mul(acc, Lift<FT, FC>::go(d_inv));
  }
  ZS_HIP(hipMemcpyAsync(b + o_lam, lam.data(), D * sizeof(TC), hipMemcpyHostToDevice, s));
  size_t g2 = ((size_t)K * hn + 255) / 256;
  if (g2 > 4096) g2 = 4096;
  TC* v = v_dev ? v_dev : (TC*)(b + o_v);
  weighted_refold_kernel<FT, FC><<<(uint32_t)g2, 256, 0, s>>>(
      in, N, K, skip, hn_log, (const TC*)(b + o_lam), v, hn);
  ZS_HIP(hipGetLastError());
  ZS_HIP(hipStreamSynchronize(s));
  if (msg0_out) memcpy(msg0_out, msg0.data(), msg0.size() * sizeof(TT));
  if (raw_out) memcpy(raw_out, raw.data(), sizeof(TC));
  *reduced_out = reduced;
  return ZS_OK;
}

template <class FT, class FC, int K>
int skip_prove_typed(zs_ctx* ctx, const zs_prove_args& a, uint32_t skip, zs_transcript* t,
                     void* msg0_out, zs_proof* out) {
  using TC = typename FC::T;
  const uint32_t hn_log = a.rounds - skip;
  const size_t hn = (size_t)1 << hn_log;
  PoolBlock vblk, blk;
  int rc = vblk.alloc(ctx, (size_t)K * hn * sizeof(TC));
  if (rc) return rc;
  TC* v = (TC*)vblk.p;
  std::vector<uint8_t> raw(sizeof(TC));
  TC reduced;
  if ((rc = skip_bind_typed<FT, FC, K>(ctx, a, skip, t, msg0_out, raw.data(), &reduced, v, blk)))
    return rc;
  hipStream_t s = ctx->stream;
  if (hn_log == 0) {
    // Every variable was skipped: the bound prism is the final value.
    std::vector<TC> fin(K);
    ZS_HIP(hipMemcpyAsync(fin.data(), v, K * sizeof(TC), hipMemcpyDeviceToHost, s));
    ZS_HIP(hipStreamSynchronize(s));
    if (out->point) memcpy(out->point, raw.data(), sizeof(TC));
    if (out->value) memcpy(out->value, &reduced, sizeof(TC));
    if (out->final_tables) memcpy(out->final_tables, fin.data(), K * sizeof(TC));
    return ZS_OK;
  }
  zs_proof tail = *out;
  if (out->point) {
    memcpy(out->point, raw.data(), sizeof(TC));
    tail.point = (uint8_t*)out->point + sizeof(TC);
  }
  return prove_tail(ctx, a, a.chal_field, ZS_MODE_PRODUCT, v, hn_log, &reduced, nullptr,
                    nullptr, t, &tail);
}

template <class FT, class FC>
int skip_bind_k(zs_ctx* ctx, const zs_prove_args& a, uint32_t skip, zs_transcript* t, void* msg0,
                void* raw, void* reduced, void* v) {
  using TC = typename FC::T;
  PoolBlock blk;
  TC red;
  int rc = ZS_OK;
  switch (a.num_tables) {
    case 1: rc = skip_bind_typed<FT, FC, 1>(ctx, a, skip, t, msg0, (uint8_t*)raw, &red, (TC*)v, blk); break;
    case 2: rc = skip_bind_typed<FT, FC, 2>(ctx, a, skip, t, msg0, (uint8_t*)raw, &red, (TC*)v, blk); break;
    case 3: rc = skip_bind_typed<FT, FC, 3>(ctx, a, skip, t, msg0, (uint8_t*)raw, &red, (TC*)v, blk); break;
    case 4: rc = skip_bind_typed<FT, FC, 4>(ctx, a, skip, t, msg0, (uint8_t*)raw, &red, (TC*)v, blk); break;
    default: return fail(ZS_EINVAL, "between 1 and 4 stacked factor tables are supported");
  }
  if (rc) return rc;
  if (reduced) memcpy(reduced, &red, sizeof(TC));
  return ZS_OK;
}

template <class FT, class FC>
int skip_dispatch_k(zs_ctx* ctx, const zs_prove_args& a, uint32_t skip, zs_transcript* t,
                    void* msg0, zs_proof* out) {
  switch (a.num_tables) {
    case 1: return skip_prove_typed<FT, FC, 1>(ctx, a, skip, t, msg0, out);
    case 2: return skip_prove_typed<FT, FC, 2>(ctx, a, skip, t, msg0, out);
    case 3: return skip_prove_typed<FT, FC, 3>(ctx, a, skip, t, msg0, out);
    case 4: return skip_prove_typed<FT, FC, 4>(ctx, a, skip, t, msg0, out);
  }
  return fail(ZS_EINVAL, "between 1 and 4 stacked factor tables are supported");
}

// -------------------------------------------------- small-value accumulators
// The accumulator pass of the small-value engine (accumulators.py:41-91,
// small_value.py:135-145). Every accumulator A_i(v, u) is a contraction of
//   T[beta] = sum_x eq(w[l0:], x) prod_q p_q(beta, x),
// beta in U_d^l0 (U_d = {inf, 0, 1, .., d-1}), x the l - l0 low variables:
// e_in x e_out of round i factor as eq(w[l0:], x) eq(w[i:l0], y). One pass
// over the input forms T: the value of table q at (beta, x) is a fixed integer
// combination of its 2^l0 prefix values at x (per variable inf -> (-1, 1),
// node t -> (1 - t, t)): the nbeta x 2^l0 matrix `cmat`, built on the host.
// Thread (column c, beta) of a block forms the base-field product for its
// beta at each column it visits, weighted by eq_hi of the column's high bits
// (lazy challenge-field sums); the low bits are fixed per thread (the grid
// stride is a multiple of 2^lo_bits), so eq_lo is applied once at the end.
// The host contracts T into the accumulators and runs the l0 rounds.
constexpr uint32_t SV_MAX_C = 8192;  // nbeta x 2^l0 matrix entries
constexpr uint32_t SV_TILE = 1024;   // K x 2^l0 x (256 / 2^blog) column values of a step

// The prime field under a table field: the accumulator pass's integer
// coefficients are stored there.
template <class F>
struct BaseOf {
  using type = F;
};
template <>
struct BaseOf<BB4> {
  using type = BB;
};
constexpr uint32_t SV_ROW = 1024;    // 64-bit words of a slot row (256 betas, 4 words)

template <class FT, class FC, int K>
__global__ __launch_bounds__(256) void sv_accum_kernel(const typename FT::T* in, size_t in_stride,
                                                       uint32_t l0, uint32_t slog, uint32_t lo_bits,
                                                       uint32_t blog, uint32_t nbeta,
                                                       const typename BaseOf<FT>::type::T* cmat,
                                                       const typename FC::T* eq_hi,
                                                       const typename FC::T* eq_lo,
                                                       uint64_t* slots) {
  using TT = typename FT::T;
  using TC = typename FC::T;
  using FB = typename BaseOf<FT>::type;
  __shared__ typename FB::T cm[SV_MAX_C];
  __shared__ TT tile[SV_TILE];  // [K][2^l0][cols] column values of the current step
  __shared__ TC red[256];
  const uint32_t tid = threadIdx.x;
  const uint32_t P = 1u << l0;
  for (uint32_t i = tid; i < nbeta * P; i += 256) cm[i] = cmat[i];
  const uint32_t beta = tid & ((1u << blog) - 1), c = tid >> blog;
  const uint32_t clog = 8u - blog, cols = 1u << clog;
  const size_t hn = (size_t)1 << slog;
  const size_t x0s = (size_t)blockIdx.x * cols;
  const bool live = beta < nbeta;
  const typename FB::T* cb = cm + (size_t)(live ? beta : 0) * P;
  const uint32_t nt = (uint32_t)K << (l0 + clog);
  WeightedSum<FC, FT> acc;
  acc.init();
  for (size_t x0 = x0s; x0 < hn; x0 += (size_t)gridDim.x * cols) {
    __syncthreads();
    for (uint32_t i = tid; i < nt; i += 256) {
      const uint32_t cc = i & (cols - 1), qb = i >> clog;
      const uint32_t q = qb >> l0, b = qb & (P - 1);
      const size_t x = x0 + cc;
      tile[i] = x < hn ? in[(size_t)q * in_stride + ((size_t)b << slog) + x] : FT::zero();
    }
    __syncthreads();
    const size_t x = x0 + c;
    if (live && x < hn) {
      TT prod = FT::one();
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const TT* col = tile + ((size_t)q << (l0 + clog)) + c;
        TT e = FT::zero();
        for (uint32_t b = 0; b < P; ++b)
          e = FT::add(e, MixMul<FT, FB>::go(col[b << clog], cb[b]));
        prod = q == 0 ? e : FT::mul(prod, e);
      }
      acc.add(eq_hi[x >> lo_bits], prod);
    }
  }
  red[tid] = FC::mul(acc.get(), eq_lo[(x0s + c) & (((size_t)1 << lo_bits) - 1)]);
  __syncthreads();
  if (tid < nbeta) {
    TC s = FC::zero();
    for (uint32_t i = tid; i < 256; i += (1u << blog)) s = FC::add(s, red[i]);
    constexpr int W = SlotOps<FC>::W;
    uint64_t* row = slots + (size_t)(blockIdx.x & (NSLOT - 1)) * SV_ROW;
#pragma unroll
    for (int c2 = 0; c2 < W; ++c2)
      __hip_atomic_fetch_add(row + tid * W + c2, (uint64_t)SlotOps<FC>::word(s, c2),
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <>
bb4 slot_total<BB4>(const uint64_t* rows, size_t row, size_t i) {
  uint32_t c[4];
  for (int w = 0; w < 4; ++w) c[w] = slot_total<BB>(rows, row, 4 * i + w);
  return bb4_make(c[0], c[1], c[2], c[3]);
}

// eq(coords, x) over {0,1}^m, coords[0] on the index MSB (poly/eq.py:92-115).
template <class F>
std::vector<typename F::T> eq_table_host(const typename F::T* coords, uint32_t m) {
  std::vector<typename F::T> t(1, F::one());
  for (uint32_t j = 0; j < m; ++j) {
    std::vector<typename F::T> nx(t.size() * 2);
    for (size_t b = 0; b < t.size(); ++b) {
      const typename F::T hi = F::mul(t[b], coords[j]);
      nx[2 * b] = F::sub(t[b], hi);
      nx[2 * b + 1] = hi;
    }
    t.swap(nx);
  }
  return t;
}

template <class FT, class FC, int K>
int sv_prove_typed(zs_ctx* ctx, const zs_prove_args& a, uint32_t l0, zs_transcript* t,
                   zs_proof* out) {
  using TT = typename FT::T;
  using TC = typename FC::T;
  const uint32_t d = K, n = a.rounds, slog = n - l0;
  uint32_t nbeta = 1;
  for (uint32_t i = 0; i < l0; ++i) nbeta *= d + 1;
  uint32_t blog = 0;
  while ((1u << blog) < nbeta) ++blog;
  const uint32_t P = 1u << l0;
  if (nbeta > 256 || (size_t)nbeta * P > SV_MAX_C ||
      ((size_t)K * P << (8 - blog)) > SV_TILE)
    return fail(ZS_EINVAL, "l_0 too large for the accumulator pass: (degree + 1)^l_0 = " +
                               std::to_string(nbeta) + " (at most 256)");
  std::vector<TC> w(n);
  for (uint32_t i = 0; i < n; ++i) w[i] = load_host_el<FC>(a.eq_point, i);
  TC claim = load_host_el<FC>(a.claim, 0);
  // cmat[beta][b]: the integer weights of the prefix values (MSB-first digits
  // of beta over U_d, MSB-first bits of b).
  using TB = typename BaseOf<FT>::type::T;
  std::vector<TB> cmat((size_t)nbeta * P);
  for (uint32_t be = 0; be < nbeta; ++be) {
    uint32_t digits[16];
    uint32_t v = be;
    for (int i = (int)l0 - 1; i >= 0; --i) {
      digits[i] = v % (d + 1);
      v /= d + 1;
    }
    for (uint32_t bb = 0; bb < P; ++bb) {
      int64_t cf = 1;
      for (uint32_t i = 0; i < l0; ++i) {
        const uint32_t bit = (bb >> (l0 - 1 - i)) & 1u;
        const int64_t tnode = (int64_t)digits[i] - 1;  // digit 0 is inf, digit t+1 node t
        const int64_t f = digits[i] == 0 ? (bit ? 1 : -1) : (bit ? tnode : 1 - tnode);
        cf *= f;
      }
      cmat[(size_t)be * P + bb] = BaseOf<FT>::type::from_int(cf);
    }
  }
  // eq(w[l0:], x) = eq_hi(x_hi) eq_lo(x_lo) with lo_bits low bits fixed per
  // thread by a power-of-two grid stride.
  const size_t hn = (size_t)1 << slog;
  const uint32_t cols = 256u >> blog;
  size_t grid = (hn + cols - 1) / cols;
  if (grid > 1024) grid = 1024;
  while (grid & (grid - 1)) grid &= grid - 1;
  uint32_t lo_bits = 0;
  while (lo_bits < slog && ((size_t)1 << (lo_bits + 1)) <= grid * cols) ++lo_bits;
  const uint32_t hi_bits = slog - lo_bits;
  std::vector<TC> eq_hi = eq_table_host<FC>(w.data() + l0, hi_bits);
  std::vector<TC> eq_lo = eq_table_host<FC>(w.data() + l0 + hi_bits, lo_bits);
  const size_t N = (size_t)1 << n;
  hipStream_t s = ctx->stream;
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off = align256(off + bytes); return o; };
  const size_t o_in = a.tables_on_device ? 0 : take((size_t)K * N * sizeof(TT));
  const size_t o_c = take(cmat.size() * sizeof(TB));
  const size_t o_hi = take(eq_hi.size() * sizeof(TC));
  const size_t o_lo = take(eq_lo.size() * sizeof(TC));
  const size_t o_slots = take((size_t)NSLOT * SV_ROW * sizeof(uint64_t));
  const size_t o_e = take((size_t)P * sizeof(TC));
  const size_t o_v = take((size_t)K * hn * sizeof(TC));
  PoolBlock blk;
  int rc = blk.alloc(ctx, off);
  if (rc) return rc;
  uint8_t* b = (uint8_t*)blk.p;
  const TT* in = a.tables_on_device ? (const TT*)a.tables : (const TT*)(b + o_in);
  if (!a.tables_on_device)
    ZS_HIP(hipMemcpyAsync(b + o_in, a.tables, (size_t)K * N * sizeof(TT), hipMemcpyHostToDevice, s));
  ZS_HIP(hipMemcpyAsync(b + o_c, cmat.data(), cmat.size() * sizeof(TB), hipMemcpyHostToDevice, s));
  ZS_HIP(hipMemcpyAsync(b + o_hi, eq_hi.data(), eq_hi.size() * sizeof(TC), hipMemcpyHostToDevice, s));
  ZS_HIP(hipMemcpyAsync(b + o_lo, eq_lo.data(), eq_lo.size() * sizeof(TC), hipMemcpyHostToDevice, s));
  ZS_HIP(hipMemsetAsync(b + o_slots, 0, (size_t)NSLOT * SV_ROW * sizeof(uint64_t), s));
  hipEvent_t ev = nullptr;
  if (ctx->profile) {
    ev = ctx->event();
    (void)hipEventRecord(ev, s);
  }
  sv_accum_kernel<FT, FC, K><<<(uint32_t)grid, 256, 0, s>>>(
      in, N, l0, slog, lo_bits, blog, nbeta, (const TB*)(b + o_c), (const TC*)(b + o_hi),
      (const TC*)(b + o_lo), (uint64_t*)(b + o_slots));
  ZS_HIP(hipGetLastError());
  if (ev) {
    hipEvent_t stop = ctx->event();
    (void)hipEventRecord(stop, s);
    ctx->records.push_back(KernelRecord{0, KREC_ENGINE, (uint32_t)grid,
                                        (uint64_t)K * N * sizeof(TT), ev, stop});
  }
  std::vector<uint64_t> rows((size_t)NSLOT * SV_ROW);
  ZS_HIP(hipMemcpyAsync(rows.data(), b + o_slots, rows.size() * sizeof(uint64_t),
                        hipMemcpyDeviceToHost, s));
  ZS_HIP(hipStreamSynchronize(s));
  std::vector<TC> T(nbeta);
  for (uint32_t be = 0; be < nbeta; ++be) T[be] = slot_total<FC>(rows.data(), SV_ROW, be);
  // Accumulators A_i[v][u] = sum_y eq(w[i:l0], y) T[beta(v, u, y)]
  // (accumulators.py:67-91 with v and y read most-significant digit first).
  std::vector<std::vector<TC>> acc(l0);
  {
    uint32_t vsize = 1;
    for (uint32_t i = 1; i <= l0; ++i) {
      acc[i - 1].assign((size_t)vsize * d, FC::zero());
      vsize *= d + 1;
    }
  }
  std::vector<std::vector<TC>> ey(l0 + 1);
  for (uint32_t i = 1; i <= l0; ++i) ey[i] = eq_table_host<FC>(w.data() + i, l0 - i);
  for (uint32_t be = 0; be < nbeta; ++be) {
    uint32_t digits[16];
    uint32_t v0 = be;
    for (int i = (int)l0 - 1; i >= 0; --i) {
      digits[i] = v0 % (d + 1);
      v0 /= d + 1;
    }
    for (uint32_t i = 1; i <= l0; ++i) {
      const uint32_t ud = digits[i - 1];
      if (ud == 2) continue;  // node 1 is not in U-hat_d
      bool cube = true;
      uint32_t y = 0;
      for (uint32_t j = i; j < l0; ++j) {
        if (digits[j] != 1 && digits[j] != 2) cube = false;
        y = 2 * y + (digits[j] == 2 ? 1u : 0u);
      }
      if (!cube) continue;
      const uint32_t u = ud > 2 ? ud - 1 : ud;
      uint32_t v = 0;
      for (uint32_t j = 0; j + 1 < i; ++j) v = (d + 1) * v + digits[j];
      TC& slot = acc[i - 1][(size_t)v * d + u];
      slot = FC::add(slot, FC::mul(ey[i][y], T[be]));
    }
  }
  // The U-hat_d wire (inf first, then 0, 2, .., d-1) and its claim reduction.
  std::vector<TC> nodes;
  nodes.push_back(FC::zero());
  for (uint32_t u = 2; u < d; ++u) nodes.push_back(FC::from_int(u));
  zs_wire uw;
  uw.kind = ZS_WIRE_DOMAIN;
  uw.num_nodes = (uint32_t)nodes.size();
  uw.has_inf = 1;
  uw.inf_pos = 0;
  uw.nodes = nodes.data();
  Ctl<FC> wc;
  memset(&wc, 0, sizeof(wc));
  if ((rc = set_ctl_wire<FC>(wc, d + 2, &uw))) return rc;
  // The accumulator rounds (small_value.py:58-108): t = R . A_i, the message
  // l_i t on U-hat_d, one transcript word a challenge, the domain's claim
  // reduction, R grown by the round's Lagrange tensor, eq mass and eq table
  // advanced.
  std::vector<TC> R(1, FC::one()), E(1, FC::one());
  TC eq_prev = FC::one();
  const size_t item = t->item_bytes;
  std::vector<TC> msgs((size_t)l0 * d);
  std::vector<uint8_t> points((size_t)l0 * sizeof(TC));
  for (uint32_t i = 0; i < l0; ++i) {
    const std::vector<TC>& A = acc[i];
    std::vector<TC> tv(d, FC::zero());
    for (size_t v = 0; v < R.size(); ++v)
      for (uint32_t u = 0; u < d; ++u) tv[u] = FC::add(tv[u], FC::mul(R[v], A[v * d + u]));
    const TC l1 = FC::mul(eq_prev, w[i]);
    const TC lz = FC::sub(eq_prev, l1);
    const TC ld = FC::sub(l1, lz);
    TC* m = msgs.data() + (size_t)i * d;
    m[0] = FC::mul(ld, tv[0]);
    for (uint32_t j = 1; j < d; ++j) m[j] = FC::mul(FC::add(lz, FC::mul(nodes[j - 1], ld)), tv[j]);
    uint8_t* raw = points.data() + (size_t)i * sizeof(TC);
    if (!transcript_hop(t, reinterpret_cast<const uint8_t*>(m), (size_t)d * sizeof(TC), item, 1,
                        raw))
      return fail(ZS_EINVAL, "transcript stream exceeds the BLAKE3 subtree stack");
    TC r;
    memcpy(&r, raw, sizeof(TC));
    r = FC::canon(r);
    // reduce_domain on U-hat_d: claim' = sum_k r^k (gamma_k claim + sum_j C[k][j] m_j).
    TC red = FC::zero(), rk = FC::one();
    for (uint32_t k = 0; k < wc.ncoef; ++k) {
      TC ck = FC::mul(wc.cclaim[k], claim);
      for (uint32_t j = 0; j < d; ++j) ck = FC::add(ck, FC::mul(wc.cmat[k][j], m[j]));
      red = FC::add(red, FC::mul(ck, rk));
      rk = FC::mul(rk, r);
    }
    claim = red;
    // Lagrange tensor over U_d at r: [prod_t (r - t), L_0(r), .., L_{d-1}(r)].
    std::vector<TC> lag(d + 1);
    lag[0] = FC::one();
    for (uint32_t tn = 0; tn < d; ++tn) lag[0] = FC::mul(lag[0], FC::sub(r, FC::from_int(tn)));
    for (uint32_t j = 0; j < d; ++j) {
      TC num = FC::one(), den = FC::one();
      for (uint32_t k = 0; k < d; ++k) {
        if (k == j) continue;
        num = FC::mul(num, FC::sub(r, FC::from_int(k)));
        den = FC::mul(den, FC::from_int((int64_t)j - (int64_t)k));
      }
      lag[j + 1] = FC::mul(num, FC::inv(den));
    }
    std::vector<TC> R2(R.size() * (d + 1));
    for (size_t v = 0; v < R.size(); ++v)
      for (uint32_t j = 0; j <= d; ++j) R2[v * (d + 1) + j] = FC::mul(R[v], lag[j]);
    R.swap(R2);
    const TC one = FC::one();
    eq_prev = FC::mul(eq_prev, FC::add(FC::mul(r, w[i]),
                                       FC::mul(FC::sub(one, r), FC::sub(one, w[i]))));
    std::vector<TC> E2(E.size() * 2);
    for (size_t bb = 0; bb < E.size(); ++bb) {
      const TC hi = FC::mul(E[bb], r);
      E2[2 * bb] = FC::sub(E[bb], hi);
      E2[2 * bb + 1] = hi;
    }
    E.swap(E2);
  }
  // Transition and tail: the factors refolded over the bound prefix, then
  // eq-poly rounds over w[l0:] from the bound eq mass; the first of them is
  // the transition round over [P_1..P_d, eq(w, .)] (small_value.py:111-128).
  ZS_HIP(hipMemcpyAsync(b + o_e, E.data(), P * sizeof(TC), hipMemcpyHostToDevice, s));
  size_t g2 = ((size_t)K * hn + 255) / 256;
  if (g2 > 4096) g2 = 4096;
  weighted_refold_kernel<FT, FC><<<(uint32_t)g2, 256, 0, s>>>(in, N, K, l0, slog,
                                                              (const TC*)(b + o_e),
                                                              (TC*)(b + o_v), hn);
  ZS_HIP(hipGetLastError());
  const uint32_t mlen = d;
  if (out->msgs) memcpy(out->msgs, msgs.data(), msgs.size() * sizeof(TC));
  if (out->point) memcpy(out->point, points.data(), points.size());
  zs_proof tail = *out;
  if (out->msgs) tail.msgs = (uint8_t*)out->msgs + (size_t)l0 * mlen * sizeof(TC);
  if (out->point) tail.point = (uint8_t*)out->point + (size_t)l0 * sizeof(TC);
  zs_prove_args ta = a;
  ta.wire = &uw;
  return prove_tail(ctx, ta, a.chal_field, ZS_MODE_EQ, b + o_v, slog, &claim, w.data() + l0,
                    &eq_prev, t, &tail);
}

template <class FT, class FC>
int sv_dispatch_k(zs_ctx* ctx, const zs_prove_args& a, uint32_t l0, zs_transcript* t,
                  zs_proof* out) {
  switch (a.num_tables) {
    case 2: return sv_prove_typed<FT, FC, 2>(ctx, a, l0, t, out);
    case 3: return sv_prove_typed<FT, FC, 3>(ctx, a, l0, t, out);
    case 4: return sv_prove_typed<FT, FC, 4>(ctx, a, l0, t, out);
  }
  return fail(ZS_EINVAL, "the small-value engine needs degree >= 2: U_1 = {inf, 0} routes no "
                         "cube point through node 1");
}

}  // namespace zs

using namespace zs;

// Structural checks first (a CPU caller sees them without a device), then the
// context.
static int skip_check(zs_ctx* ctx, const zs_prove_args* args, uint32_t skip_rounds,
                      zs_transcript* transcript) {
  uint32_t npts = 0;
  int rc = validate_args(*args, transcript, &npts);
  if (rc) return rc;
  if (args->mode != ZS_MODE_PRODUCT)
    return fail(ZS_EINVAL, "the univariate skip takes the product summand");
  if (args->refold_rounds) return fail(ZS_EINVAL, "refold_rounds must be 0 with the skip");
  if (skip_rounds < 1 || skip_rounds > args->rounds)
    return fail(ZS_EINVAL, "skip_rounds must be in [1, " + std::to_string(args->rounds) + "]");
  if (args->table_field == ZS_F_BABYBEAR4)
    return fail(ZS_EINVAL, "the skip's round 0 is base-field work: tables must be BabyBear or "
                           "Goldilocks");
  if (transcript->field != args->table_field)
    return fail(ZS_EINVAL, "the round-0 message is base-field: the transcript must be over the "
                           "tables' field");
  if (!ctx) return fail(ZS_EINVAL, "null context");
  return set_device(ctx);
}

extern "C" int zs_prove_univariate_skip(zs_ctx* ctx, const zs_prove_args* args,
                                        uint32_t skip_rounds, zs_transcript* transcript,
                                        void* msg0, zs_proof* out) {
  if (!args || !transcript || !out) return fail(ZS_EINVAL, "null argument");
  int rc = skip_check(ctx, args, skip_rounds, transcript);
  if (rc) return rc;
  if (args->table_field == ZS_F_BABYBEAR && args->chal_field == ZS_F_BABYBEAR)
    return skip_dispatch_k<BB, BB>(ctx, *args, skip_rounds, transcript, msg0, out);
  if (args->table_field == ZS_F_BABYBEAR && args->chal_field == ZS_F_BABYBEAR4)
    return skip_dispatch_k<BB, BB4>(ctx, *args, skip_rounds, transcript, msg0, out);
  if (args->table_field == ZS_F_GOLDILOCKS && args->chal_field == ZS_F_GOLDILOCKS)
    return skip_dispatch_k<GL, GL>(ctx, *args, skip_rounds, transcript, msg0, out);
  return fail(ZS_EINVAL, "unsupported field combination");
}

extern "C" int zs_univariate_skip_round0(zs_ctx* ctx, const zs_prove_args* args,
                                         uint32_t skip_rounds, zs_transcript* transcript,
                                         void* msg0, void* chal, void* reduced,
                                         void* bound_tables) {
  if (!args || !transcript || !bound_tables) return fail(ZS_EINVAL, "null argument");
  int rc = skip_check(ctx, args, skip_rounds, transcript);
  if (rc) return rc;
  if (args->table_field == ZS_F_BABYBEAR && args->chal_field == ZS_F_BABYBEAR)
    return skip_bind_k<BB, BB>(ctx, *args, skip_rounds, transcript, msg0, chal, reduced,
                               bound_tables);
  if (args->table_field == ZS_F_BABYBEAR && args->chal_field == ZS_F_BABYBEAR4)
    return skip_bind_k<BB, BB4>(ctx, *args, skip_rounds, transcript, msg0, chal, reduced,
                                bound_tables);
  if (args->table_field == ZS_F_GOLDILOCKS && args->chal_field == ZS_F_GOLDILOCKS)
    return skip_bind_k<GL, GL>(ctx, *args, skip_rounds, transcript, msg0, chal, reduced,
                               bound_tables);
  return fail(ZS_EINVAL, "unsupported field combination");
}

extern "C" int zs_prove_small_value(zs_ctx* ctx, const zs_prove_args* args, uint32_t l0,
                                    zs_transcript* transcript, zs_proof* out) {
  if (!args || !transcript || !out) return fail(ZS_EINVAL, "null argument");
  uint32_t npts = 0;
  int rc = validate_args(*args, transcript, &npts);
  if (rc) return rc;
  if (args->mode != ZS_MODE_EQ) return fail(ZS_EINVAL, "the small-value engine is eq-weighted "
                                                       "(ZS_MODE_EQ)");
  if (args->wire) return fail(ZS_EINVAL, "the small-value engine sends compressed U-hat "
                                         "messages: its wire form is fixed (wire must be null)");
  if (args->eq_mass) return fail(ZS_EINVAL, "the small-value engine starts with nothing bound");
  const uint32_t n = args->rounds;
  if (n < 2) return fail(ZS_EINVAL, "the small-value engine needs at least 2 rounds");
  if (l0 < 1 || l0 > n - n / 2)
    return fail(ZS_EINVAL, "l_0 must be in [1, " + std::to_string(n - n / 2) +
                               "] (the small-value rounds fit before the out-half); got l_0=" +
                               std::to_string(l0) + " for l=" + std::to_string(n));
  if (transcript->item_bytes != field_bytes(args->chal_field))
    return fail(ZS_EINVAL, "the small-value rounds squeeze one transcript word a challenge: the "
                           "transcript must be over the challenge field");
  if (args->num_tables < 2)
    return fail(ZS_EINVAL, "the small-value engine needs degree >= 2: U_1 = {inf, 0} routes no "
                           "cube point through node 1");
  if (!ctx) return fail(ZS_EINVAL, "null context");
  if ((rc = set_device(ctx))) return rc;
  if (args->table_field == ZS_F_BABYBEAR && args->chal_field == ZS_F_BABYBEAR)
    return sv_dispatch_k<BB, BB>(ctx, *args, l0, transcript, out);
  if (args->table_field == ZS_F_BABYBEAR && args->chal_field == ZS_F_BABYBEAR4)
    return sv_dispatch_k<BB, BB4>(ctx, *args, l0, transcript, out);
  if (args->table_field == ZS_F_BABYBEAR4 && args->chal_field == ZS_F_BABYBEAR4)
    return sv_dispatch_k<BB4, BB4>(ctx, *args, l0, transcript, out);
  if (args->table_field == ZS_F_GOLDILOCKS && args->chal_field == ZS_F_GOLDILOCKS)
    return sv_dispatch_k<GL, GL>(ctx, *args, l0, transcript, out);
  return fail(ZS_EINVAL, "unsupported field combination");
}
