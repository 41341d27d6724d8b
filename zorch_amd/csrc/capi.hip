// C ABI of the prover: context, host transcript, whole-stage and per-round
// proving, batched proving and the host verifier step.
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/zorch_amd.h"
#include "blake3.h"
#include "field.h"
#include "runtime.h"
#include "sha256.h"

namespace zs {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}
int hip_fail(hipError_t e, const char* what) {
  return fail(e == hipErrorOutOfMemory ? ZS_ENOMEM : ZS_EHIP,
              std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------------- pool
DevicePool::~DevicePool() {
  trim_locked();
  for (auto& kv : sizes_) (void)hipFree(kv.first);
}

void DevicePool::trim() {
  std::lock_guard<std::mutex> g(mu_);
  trim_locked();
}

void DevicePool::trim_locked() {
  for (auto& kv : free_) {
    (void)hipFree(kv.second);
    sizes_.erase(kv.second);
  }
  free_.clear();
  cached_ = 0;
}

int DevicePool::alloc(size_t bytes, void** out) {
  bytes = (bytes + 4095) & ~(size_t)4095;
  std::lock_guard<std::mutex> g(mu_);
  auto it = free_.find(bytes);
  if (it != free_.end()) {
    *out = it->second;
    free_.erase(it);
    cached_ -= bytes;
    live_ += bytes;
    return ZS_OK;
  }
  if (live_ + bytes > budget_) {
    return fail(ZS_EBUDGET, "device memory budget exceeded: " + std::to_string(live_ + bytes) +
                                " > " + std::to_string(budget_) + " bytes");
  }
  if (live_ + cached_ + bytes > budget_) trim_locked();
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    trim_locked();
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  }
  sizes_[p] = bytes;
  live_ += bytes;
  *out = p;
  return ZS_OK;
}

void DevicePool::release(void* p) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = sizes_.find(p);
  if (it == sizes_.end()) return;
  live_ -= it->second;
  cached_ += it->second;
  free_.emplace(it->second, p);
}

HostPool::~HostPool() {
  for (auto& kv : sizes_) (void)hipHostFree(kv.first);
}

int HostPool::alloc(size_t bytes, void** out) {
  bytes = (bytes + 4095) & ~(size_t)4095;
  std::lock_guard<std::mutex> g(mu_);
  auto it = free_.find(bytes);
  if (it != free_.end()) {
    *out = it->second;
    free_.erase(it);
    return ZS_OK;
  }
  void* p = nullptr;
  hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
  if (e != hipSuccess) return hip_fail(e, "hipHostMalloc");
  sizes_[p] = bytes;
  *out = p;
  return ZS_OK;
}

void HostPool::release(void* p) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = sizes_.find(p);
  if (it == sizes_.end()) return;
  free_.emplace(it->second, p);
}

// ---------------------------------------------------------------- helpers
uint32_t field_bytes(uint32_t field) {
  switch (field) {
    case ZS_F_BABYBEAR: return 4;
    case ZS_F_BABYBEAR4: return 16;
    case ZS_F_GOLDILOCKS: return 8;
    default: return 0;
  }
}

static bool known_kind(uint32_t kind) { return kind == ZS_TR_SHA256 || kind == ZS_TR_BLAKE3; }

// One Fiat-Shamir hop (observe under slice framing, sample `limbs` words) on
// either transcript kind; false if a BLAKE3 stream outgrows its subtree stack.
bool transcript_hop(zs_transcript* t, const uint8_t* msg, size_t len, size_t item, size_t limbs,
                    uint8_t* out) {
  if (t->kind == ZS_TR_BLAKE3)
    return b3_fs_hop(*reinterpret_cast<Blake3State*>(&t->blake3), msg, len, item, limbs, out);
  sha256_fs_hop(*reinterpret_cast<Sha256State*>(&t->sha), msg, len, item, limbs, out);
  return true;
}

// Base prime field of a field id.
uint32_t base_of(uint32_t f) { return f == ZS_F_BABYBEAR4 ? ZS_F_BABYBEAR : f; }

int validate_args(const zs_prove_args& a, const zs_transcript* t, uint32_t* npts,
                  uint32_t* mlen) {
  if (!field_bytes(a.table_field) || !field_bytes(a.chal_field))
    return fail(ZS_EINVAL, "unknown field id");
  const bool ok_pair = (a.table_field == a.chal_field) ||
                       (a.table_field == ZS_F_BABYBEAR && a.chal_field == ZS_F_BABYBEAR4);
  if (!ok_pair) return fail(ZS_EINVAL, "table field does not embed in the challenge field");
  if (a.mode != ZS_MODE_PRODUCT && a.mode != ZS_MODE_EQ && a.mode != ZS_MODE_LOGUP)
    return fail(ZS_EINVAL, "unknown mode");
  if (a.mode == ZS_MODE_LOGUP) {
    // The LogUp summand (zorch/logup_gkr/prover.py:82-104): five factors,
    // degree 3, one field, natural values on the wire.
    if (a.num_tables != 5 || a.degree != 3)
      return fail(ZS_EINVAL, "LogUp combine needs 5 factors [eq, n0, d1, n1, d0] of degree 3");
    if (a.table_field != a.chal_field)
      return fail(ZS_EINVAL, "LogUp tables must be in the challenge field");
    if (!a.lambda) return fail(ZS_EINVAL, "a LogUp stage needs its batching challenge lambda");
    if (a.eq_point || a.eq_mass) return fail(ZS_EINVAL, "a LogUp stage folds its eq table itself");
    if (a.wire && a.wire->kind != ZS_WIRE_VALUES)
      return fail(ZS_EINVAL, "a LogUp round sends its natural values");
    if (a.refold_rounds) return fail(ZS_EINVAL, "refold_rounds must be 0 for a LogUp stage");
  } else {
    if (a.degree < 1) return fail(ZS_EINVAL, "degree must be >= 1");
    if (a.num_tables < 1 || a.num_tables > 4)
      return fail(ZS_EINVAL, "between 1 and 4 stacked factor tables are supported");
    if (a.degree != a.num_tables)
      return fail(ZS_EINVAL,
                  "the product summand's degree must equal the number of stacked tables");
  }
  // Table lengths of a round (2^(n - r) entries) are 32-bit in the kernels.
  if (a.rounds < 1 || a.rounds > ZS_MAX_ROUNDS)
    return fail(ZS_EINVAL, "rounds must be in [1, " + std::to_string(ZS_MAX_ROUNDS) + "]");
  if (!a.tables || !a.claim) return fail(ZS_EINVAL, "null tables or claim");
  if (a.mode == ZS_MODE_EQ && !a.eq_point) return fail(ZS_EINVAL, "eq mode needs an equality point");
  if (a.refold_rounds > a.rounds / 2)
    return fail(ZS_EINVAL, "refold_rounds must be at most rounds / 2");
  if (a.refold_rounds && a.mode != ZS_MODE_PRODUCT)
    return fail(ZS_EINVAL, "refold rounds (the sqrt-space engine) take the product summand");
  if (t) {
    if (!known_kind(t->kind)) return fail(ZS_EINVAL, "unknown transcript kind");
    const uint32_t tf = t->field;
    if (!(tf == a.chal_field || tf == base_of(a.chal_field)))
      return fail(ZS_EINVAL, "challenge field does not tile the transcript field");
    if (t->item_bytes != field_bytes(tf)) return fail(ZS_EINVAL, "transcript item size mismatch");
  }
  *npts = a.degree + 1 + (a.mode == ZS_MODE_EQ ? 1 : 0);
  const uint32_t m = wire_width(a.wire, *npts, a.mode == ZS_MODE_PRODUCT && a.num_tables == 2);
  if (!m) return ZS_EINVAL;
  if ((size_t)m * field_bytes(a.chal_field) > 96)
    return fail(ZS_EINVAL, "a round message is at most 96 bytes");
  if (mlen) *mlen = m;
  return ZS_OK;
}

// Dispatch to the unit of a field pair and factor count (nullptr: unsupported).
#define ZS_DISPATCH(WHAT)                                                              \
  if (k < 1 || k > 4) return nullptr;                                                  \
  if (tf == ZS_F_BABYBEAR && cf == ZS_F_BABYBEAR) {                                    \
    static decltype(&WHAT##_bb_k1) const f[4] = {WHAT##_bb_k1, WHAT##_bb_k2,            \
                                                 WHAT##_bb_k3, WHAT##_bb_k4};           \
    return f[k - 1](npts, eq);                                                         \
  }                                                                                    \
  if (tf == ZS_F_BABYBEAR && cf == ZS_F_BABYBEAR4) {                                   \
    static decltype(&WHAT##_bb_bb4_k1) const f[4] = {WHAT##_bb_bb4_k1, WHAT##_bb_bb4_k2, \
                                                     WHAT##_bb_bb4_k3, WHAT##_bb_bb4_k4}; \
    return f[k - 1](npts, eq);                                                         \
  }                                                                                    \
  if (tf == ZS_F_BABYBEAR4 && cf == ZS_F_BABYBEAR4) {                                  \
    static decltype(&WHAT##_bb4_k1) const f[4] = {WHAT##_bb4_k1, WHAT##_bb4_k2,         \
                                                  WHAT##_bb4_k3, WHAT##_bb4_k4};        \
    return f[k - 1](npts, eq);                                                         \
  }                                                                                    \
  if (tf == ZS_F_GOLDILOCKS && cf == ZS_F_GOLDILOCKS) {                                \
    static decltype(&WHAT##_gl_k1) const f[4] = {WHAT##_gl_k1, WHAT##_gl_k2,            \
                                                 WHAT##_gl_k3, WHAT##_gl_k4};           \
    return f[k - 1](npts, eq);                                                         \
  }                                                                                    \
  return nullptr;

BatchGroup* make_batch_group(uint32_t tf, uint32_t cf, uint32_t k, uint32_t npts, bool eq) {
  ZS_DISPATCH(make_batch_group)
}

ServeKind* make_serve_kind(uint32_t tf, uint32_t cf, uint32_t k, uint32_t npts, bool eq) {
  ZS_DISPATCH(make_serve_kind)
}

Instance* make_instance(uint32_t tf, uint32_t cf, uint32_t k, uint32_t npts, bool eq) {
  ZS_DISPATCH(make_instance)
}
#undef ZS_DISPATCH

int set_device(zs_ctx* ctx) {
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  return ZS_OK;
}

// ------------------------------------------------------------- verifier
template <class F>
static int verify_round_typed(zs_transcript* t, const void* claim_p, const void* msg_p,
                              uint32_t npts, void* chal_out, void* reduced_out, int* ok) {
  using T = typename F::T;
  std::vector<T> msg(npts);
  memcpy(msg.data(), msg_p, npts * sizeof(T));
  T claim;
  memcpy(&claim, claim_p, sizeof(T));
  claim = F::canon(claim);
  for (auto& m : msg) m = F::canon(m);
  const size_t item = t->item_bytes;
  const size_t limbs = sizeof(T) / item;
  std::vector<uint8_t> raw(sizeof(T));
  if (!transcript_hop(t, reinterpret_cast<const uint8_t*>(msg_p), npts * sizeof(T), item, limbs,
                      raw.data()))
    return fail(ZS_EINVAL, "transcript stream exceeds the BLAKE3 subtree stack");
  if (chal_out) memcpy(chal_out, raw.data(), sizeof(T));
  T r;
  memcpy(&r, raw.data(), sizeof(T));
  r = F::canon(r);
  // Lagrange over {0..npts-1} at r (reference poly/univariate.py:36-65).
  T red = F::zero();
  for (uint32_t k = 0; k < npts; ++k) {
    T num = F::one(), den = F::one();
    for (uint32_t j = 0; j < npts; ++j) {
      if (j == k) continue;
      num = F::mul(num, F::sub(r, F::from_int(j)));
      den = F::mul(den, F::from_int((int64_t)k - (int64_t)j));
    }
    red = F::add(red, F::mul(msg[k], F::mul(num, F::inv(den))));
  }
  if (reduced_out) memcpy(reduced_out, &red, sizeof(T));
  *ok = F::eq(claim, F::add(msg[0], msg[1])) ? 1 : 0;
  return ZS_OK;
}

// One field operation per thread (zs_field_op).
template <class F>
__global__ void field_op_kernel(uint32_t op, const typename F::T* a, const typename F::T* b,
                                typename F::T* out, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const typename F::T x = F::canon(a[i]), y = F::canon(b[i]);
  out[i] = op == ZS_OP_ADD ? F::add(x, y) : op == ZS_OP_SUB ? F::sub(x, y) : F::mul(x, y);
}

template <class F>
static int field_op_typed(zs_ctx* ctx, uint32_t op, const void* a, const void* b, void* out,
                          size_t count) {
  using T = typename F::T;
  const size_t bytes = count * sizeof(T);
  void* d = nullptr;
  int rc = ctx->pool->alloc(3 * bytes, &d);
  if (rc) return rc;
  T* da = (T*)d;
  T* db = da + count;
  T* dout = db + count;
  hipStream_t s = ctx->stream;
  hipError_t e = hipMemcpyAsync(da, a, bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(db, b, bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    field_op_kernel<F><<<(uint32_t)((count + 255) / 256), 256, 0, s>>>(op, da, db, dout, count);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  ctx->pool->release(d);
  return e == hipSuccess ? ZS_OK : hip_fail(e, "zs_field_op");
}

// Streaming read of n 16-byte words: four independent loads in flight per
// thread and step; the xor of everything read is stored only when it equals
// a constant, which keeps the loads without writing a byte in practice.
__global__ __launch_bounds__(256) void stream_read_kernel(const uint4* __restrict__ p, size_t n,
                                                          uint32_t* sink) {
  const size_t stride = (size_t)gridDim.x * 256;
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  uint32_t acc = 0;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const uint4 a = p[i], b = p[i + stride], c = p[i + 2 * stride], d = p[i + 3 * stride];
    acc ^= a.x ^ a.y ^ a.z ^ a.w ^ b.x ^ b.y ^ b.z ^ b.w ^ c.x ^ c.y ^ c.z ^ c.w ^ d.x ^ d.y ^
           d.z ^ d.w;
  }
  for (; i < n; i += stride) {
    const uint4 a = p[i];
    acc ^= a.x ^ a.y ^ a.z ^ a.w;
  }
  if (acc == 0x9e3779b9u) sink[blockIdx.x] = acc;
}

// Streaming copy of n 16-byte words.
__global__ __launch_bounds__(256) void stream_copy_kernel(const uint4* __restrict__ src,
                                                          uint4* __restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace zs

using namespace zs;

struct zs_state {
  zs_ctx* ctx = nullptr;
  std::unique_ptr<Instance> inst;
  uint32_t bound = 0;
};

extern "C" {

const char* zs_last_error(void) { return g_last_error.c_str(); }
const char* zs_version(void) { return "zorch_amd 0.1.0 gfx950"; }

int zs_device_count(int* count) {
  if (!count) return fail(ZS_EINVAL, "null out");
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    return hip_fail(e, "hipGetDeviceCount");
  }
  return ZS_OK;
}

int zs_ctx_create(int device, uint64_t budget, zs_ctx** out) {
  if (!out) return fail(ZS_EINVAL, "null out");
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return fail(ZS_ENODEV, "no HIP device available");
  if (device < 0 || device >= count) return fail(ZS_EINVAL, "device index out of range");
  ZS_HIP(hipSetDevice(device));
  if (budget == 0) {
    size_t free_b = 0, total_b = 0;
    ZS_HIP(hipMemGetInfo(&free_b, &total_b));
    budget = (uint64_t)(free_b * 0.9);
  }
  // The no-allocation entry point's scratch array, looked up once per device
  // with the device current.
  int rc = noalloc_scratch_init(device);
  if (rc) return rc;
  auto* ctx = new zs_ctx();
  ctx->device = device;
  ctx->pool = new DevicePool(budget);
  ctx->host_pool = new HostPool();
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx->host_pool;
    delete ctx->pool;
    delete ctx;
    return hip_fail(e, "hipStreamCreate");
  }
  *out = ctx;
  return ZS_OK;
}

void zs_ctx_destroy(zs_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& r : ctx->records) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  for (auto e : ctx->spare_events) (void)hipEventDestroy(e);
  if (ctx->trace) (void)hipFree(ctx->trace);
  delete ctx->host_pool;
  delete ctx->pool;
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int zs_ctx_reserved_bytes(zs_ctx* ctx, uint64_t* out) {
  if (!ctx || !out) return fail(ZS_EINVAL, "null argument");
  *out = ctx->pool->held();
  return ZS_OK;
}

int zs_device_alloc(zs_ctx* ctx, size_t bytes, void** dptr) {
  if (!ctx || !dptr || bytes == 0) return fail(ZS_EINVAL, "bad device allocation request");
  ZS_HIP(hipSetDevice(ctx->device));
  return ctx->pool->alloc(bytes, dptr);
}

int zs_device_free(zs_ctx* ctx, void* dptr) {
  if (!ctx || !dptr) return fail(ZS_EINVAL, "null argument");
  ZS_HIP(hipSetDevice(ctx->device));
  ZS_HIP(hipStreamSynchronize(ctx->stream));
  ctx->pool->release(dptr);
  return ZS_OK;
}

int zs_memcpy_htod(zs_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return fail(ZS_EINVAL, "null argument");
  ZS_HIP(hipSetDevice(ctx->device));
  ZS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  ZS_HIP(hipStreamSynchronize(ctx->stream));
  return ZS_OK;
}

int zs_memcpy_dtoh(zs_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return fail(ZS_EINVAL, "null argument");
  ZS_HIP(hipSetDevice(ctx->device));
  ZS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  ZS_HIP(hipStreamSynchronize(ctx->stream));
  return ZS_OK;
}

int zs_memcpy_dtod(zs_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return fail(ZS_EINVAL, "null argument");
  ZS_HIP(hipSetDevice(ctx->device));
  ZS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  ZS_HIP(hipStreamSynchronize(ctx->stream));
  return ZS_OK;
}

int zs_ctx_get_tune(zs_ctx* ctx, zs_tune* out) {
  if (!ctx || !out) return fail(ZS_EINVAL, "null argument");
  *out = zs_tune{ctx->tune.tail_pairs, ctx->tune.max_grid, ctx->tune.pairs_per_thread,
                 ctx->tune.chain_pairs, ctx->tune.chain_grid, ctx->tune.step_rounds,
                 ctx->tune.lookahead, {0}};
  return ZS_OK;
}

int zs_ctx_set_tune(zs_ctx* ctx, const zs_tune* t) {
  if (!ctx || !t) return fail(ZS_EINVAL, "null argument");
  const uint32_t g = t->max_grid;
  if (g == 0 || g > zs::TUNE_MAX_GRID || (g & (g - 1)))
    return fail(ZS_EINVAL, "max_grid must be a power of two in [1, 4096]");
  if (t->pairs_per_thread == 0 || t->pairs_per_thread > (1u << 20))
    return fail(ZS_EINVAL, "pairs_per_thread must be in [1, 2^20]");
  const uint32_t cg = t->chain_grid;
  if (cg == 0 || cg > zs::TUNE_MAX_GRID || (cg & (cg - 1)))
    return fail(ZS_EINVAL, "chain_grid must be a power of two in [1, 4096]");
  if (t->lookahead > 1) return fail(ZS_EINVAL, "lookahead must be 0 or 1");
  ctx->tune.tail_pairs = t->tail_pairs;
  ctx->tune.max_grid = g;
  ctx->tune.pairs_per_thread = t->pairs_per_thread;
  ctx->tune.chain_pairs = t->chain_pairs;
  ctx->tune.chain_grid = t->chain_grid;
  ctx->tune.step_rounds = t->step_rounds;
  ctx->tune.lookahead = t->lookahead;
  return ZS_OK;
}

int zs_ctx_profile(zs_ctx* ctx, int enable) {
  if (!ctx) return fail(ZS_EINVAL, "null argument");
  ZS_HIP(hipSetDevice(ctx->device));
  ZS_HIP(hipStreamSynchronize(ctx->stream));
  for (auto& r : ctx->records) {
    ctx->spare_events.push_back(r.start);
    ctx->spare_events.push_back(r.stop);
  }
  ctx->records.clear();
  ctx->profile = enable != 0;
  return ZS_OK;
}

static const size_t kTraceWords = 4096;

int zs_ctx_trace(zs_ctx* ctx, int enable) {
  if (!ctx) return fail(ZS_EINVAL, "null argument");
  ZS_HIP(hipSetDevice(ctx->device));
  ZS_HIP(hipStreamSynchronize(ctx->stream));
  if (enable && !ctx->trace) {
    ZS_HIP(hipMalloc((void**)&ctx->trace, kTraceWords * sizeof(uint64_t)));
  }
  if (!enable && ctx->trace) {
    ZS_HIP(hipFree(ctx->trace));
    ctx->trace = nullptr;
  }
  if (ctx->trace) ZS_HIP(hipMemset(ctx->trace, 0, kTraceWords * sizeof(uint64_t)));
  return ZS_OK;
}

int zs_ctx_trace_read(zs_ctx* ctx, uint64_t* pairs, size_t cap, size_t* count,
                      uint64_t* clock_khz) {
  if (!ctx || !count) return fail(ZS_EINVAL, "null argument");
  *count = 0;
  if (!ctx->trace) return ZS_OK;
  ZS_HIP(hipSetDevice(ctx->device));
  ZS_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<uint64_t> buf(kTraceWords);
  ZS_HIP(hipMemcpy(buf.data(), ctx->trace, kTraceWords * sizeof(uint64_t), hipMemcpyDeviceToHost));
  size_t n = buf[0] < (kTraceWords - 1) / 2 ? (size_t)buf[0] : (kTraceWords - 1) / 2;
  if (n > cap) n = cap;
  if (pairs) memcpy(pairs, buf.data() + 1, n * 2 * sizeof(uint64_t));
  *count = n;
  if (clock_khz) {
    int rate = 0;
    ZS_HIP(hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, ctx->device));
    *clock_khz = (uint64_t)rate;
  }
  ZS_HIP(hipMemset(ctx->trace, 0, kTraceWords * sizeof(uint64_t)));
  return ZS_OK;
}

int zs_ctx_profile_read(zs_ctx* ctx, zs_kernel_record* out, uint32_t cap, uint32_t* count) {
  if (!ctx || !count || (cap && !out)) return fail(ZS_EINVAL, "null argument");
  ZS_HIP(hipSetDevice(ctx->device));
  ZS_HIP(hipStreamSynchronize(ctx->stream));
  uint32_t n = 0;
  for (auto& r : ctx->records) {
    if (n < cap) {
      float ms = 0.f;
      ZS_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
      out[n] = zs_kernel_record{r.round, r.kind, r.grid, 0, r.bytes, (double)ms};
    }
    ++n;
    ctx->spare_events.push_back(r.start);
    ctx->spare_events.push_back(r.stop);
  }
  ctx->records.clear();
  *count = n < cap ? n : cap;
  return ZS_OK;
}

// ------------------------------------------------------------- transcript
int zs_transcript_new(uint32_t kind, uint32_t field, const uint8_t* domain, size_t len,
                      zs_transcript* out) {
  if (!out || (len && !domain)) return fail(ZS_EINVAL, "null argument");
  if (!known_kind(kind)) return fail(ZS_EINVAL, "unknown transcript kind");
  if (!field_bytes(field)) return fail(ZS_EINVAL, "unknown field id");
  memset(out, 0, sizeof(*out));
  out->kind = kind;
  out->field = field;
  out->item_bytes = field_bytes(field);
  uint8_t head[9];
  head[0] = OP_DOMAIN;
  for (int i = 0; i < 8; ++i) head[1 + i] = (uint8_t)((uint64_t)len >> (8 * i));
  if (kind == ZS_TR_BLAKE3) {
    Blake3State* s = reinterpret_cast<Blake3State*>(&out->blake3);
    b3_init(*s);
    if (!b3_absorb(*s, head, 9) || !b3_absorb(*s, domain, len))
      return fail(ZS_EINVAL, "transcript stream exceeds the BLAKE3 subtree stack");
    return ZS_OK;
  }
  Sha256State* s = reinterpret_cast<Sha256State*>(&out->sha);
  sha256_init(*s);
  sha256_absorb(*s, head, 9);
  sha256_absorb(*s, domain, len);
  return ZS_OK;
}

int zs_transcript_absorb(zs_transcript* t, const uint8_t* data, size_t len) {
  if (!t || (len && !data)) return fail(ZS_EINVAL, "null argument");
  if (!known_kind(t->kind)) return fail(ZS_EINVAL, "unknown transcript kind");
  if (t->kind == ZS_TR_BLAKE3) {
    if (!b3_absorb(*reinterpret_cast<Blake3State*>(&t->blake3), data, len))
      return fail(ZS_EINVAL, "transcript stream exceeds the BLAKE3 subtree stack");
    return ZS_OK;
  }
  sha256_absorb(*reinterpret_cast<Sha256State*>(&t->sha), data, len);
  return ZS_OK;
}

int zs_transcript_squeeze(zs_transcript* t, size_t nbytes, uint8_t* out) {
  if (!t || (nbytes && !out)) return fail(ZS_EINVAL, "null argument");
  if (!known_kind(t->kind)) return fail(ZS_EINVAL, "unknown transcript kind");
  if (t->kind == ZS_TR_BLAKE3) {
    Blake3State* s = reinterpret_cast<Blake3State*>(&t->blake3);
    b3_finalize(*s, out, nbytes);
    if (!b3_absorb(*s, out, nbytes))
      return fail(ZS_EINVAL, "transcript stream exceeds the BLAKE3 subtree stack");
    return ZS_OK;
  }
  sha256_squeeze(*reinterpret_cast<Sha256State*>(&t->sha), out, nbytes);
  return ZS_OK;
}

int zs_transcript_digest(const zs_transcript* t, uint8_t out[32]) {
  if (!t || !out) return fail(ZS_EINVAL, "null argument");
  if (!known_kind(t->kind)) return fail(ZS_EINVAL, "unknown transcript kind");
  if (t->kind == ZS_TR_BLAKE3) {
    b3_finalize(*reinterpret_cast<const Blake3State*>(&t->blake3), out, 32);
    return ZS_OK;
  }
  sha256_digest_suffix(*reinterpret_cast<const Sha256State*>(&t->sha), nullptr, 0, out);
  return ZS_OK;
}

// ------------------------------------------------------------------ prove
int zs_prove_check(const zs_prove_args* args, const zs_transcript* transcript,
                   uint32_t* msg_len) {
  if (!args || !msg_len) return fail(ZS_EINVAL, "null argument");
  uint32_t npts = 0;
  return validate_args(*args, transcript, &npts, msg_len);
}

int zs_prove(zs_ctx* ctx, const zs_prove_args* args, zs_transcript* transcript, zs_proof* out) {
  if (!ctx || !args || !transcript || !out) return fail(ZS_EINVAL, "null argument");
  uint32_t npts = 0;
  int rc = validate_args(*args, transcript, &npts);
  if (rc) return rc;
  if ((rc = set_device(ctx))) return rc;
  std::unique_ptr<Instance> inst(
      args->mode == ZS_MODE_LOGUP
          ? make_logup_instance(args->chal_field)
          : make_instance(args->table_field, args->chal_field, args->num_tables, npts,
                          args->mode == ZS_MODE_EQ));
  if (!inst) return fail(ZS_EINVAL, "unsupported field / degree combination");
  if ((rc = inst->setup(ctx, *args, *transcript))) return rc;
  if ((rc = inst->enqueue(0, args->rounds, ctx->stream))) return rc;
  if ((rc = inst->download(ctx->stream))) return rc;
  ZS_HIP(hipStreamSynchronize(ctx->stream));
  return inst->collect(out, transcript);
}

int zs_logup_transition(zs_ctx* ctx, uint32_t field, const void* layer, size_t width,
                        void* next) {
  if (!ctx || !layer || !next) return fail(ZS_EINVAL, "null argument");
  if (!field_bytes(field)) return fail(ZS_EINVAL, "unknown field");
  if (width < 2 || (width & (width - 1)))
    return fail(ZS_EINVAL, "a layer's width must be a power of two of at least 2");
  int rc;
  if ((rc = set_device(ctx))) return rc;
  return logup_transition(ctx, field, layer, width, next);
}

int zs_logup_eq(zs_ctx* ctx, uint32_t field, const void* point, uint32_t n, void* out) {
  if (!ctx || (n && !point) || !out) return fail(ZS_EINVAL, "null argument");
  if (!field_bytes(field)) return fail(ZS_EINVAL, "unknown field");
  if (n > ZS_MAX_ROUNDS) return fail(ZS_EINVAL, "too many coordinates");
  int rc;
  if ((rc = set_device(ctx))) return rc;
  return logup_eq(ctx, field, point, n, out);
}

int zs_jagged_transition(zs_ctx* ctx, uint32_t field, uint32_t num_segments,
                         const uint32_t* row_counts, const uint32_t* out_row_counts,
                         const void* planes, size_t width, void* next, size_t out_width) {
  if (!ctx || !row_counts || !out_row_counts || !planes || !next)
    return fail(ZS_EINVAL, "null argument");
  if (!field_bytes(field)) return fail(ZS_EINVAL, "unknown field");
  if (num_segments == 0) return fail(ZS_EINVAL, "no segments");
  uint64_t live = 0, out_live = 0;
  for (uint32_t s = 0; s < num_segments; ++s) {
    if (out_row_counts[s] < (row_counts[s] + 1) / 2)
      return fail(ZS_EINVAL, "out row count " + std::to_string(out_row_counts[s]) +
                                 " truncates a segment of " + std::to_string(row_counts[s]) +
                                 " rows");
    live += row_counts[s];
    out_live += out_row_counts[s];
  }
  if (live > width) return fail(ZS_EINVAL, "the row counts exceed the layer's width");
  if (out_live > out_width || out_width > 0xffffffffull)
    return fail(ZS_EINVAL, "out_width cannot hold the schedule's live rows");
  int rc;
  if ((rc = set_device(ctx))) return rc;
  return jagged_transition(ctx, field, num_segments, row_counts, out_row_counts, planes, width,
                           next, out_width);
}

int zs_prove_jagged_layer(zs_ctx* ctx, const zs_jagged_args* args, zs_transcript* transcript,
                          zs_jagged_proof* out) {
  if (!ctx || !args || !transcript || !out) return fail(ZS_EINVAL, "null argument");
  const zs_jagged_args& a = *args;
  if (!field_bytes(a.field)) return fail(ZS_EINVAL, "unknown field");
  if (!a.row_counts || !a.planes || !a.eval_point || !a.lambda || !a.claim)
    return fail(ZS_EINVAL, "null argument");
  if (a.num_segments == 0 || (a.num_segments & (a.num_segments - 1)))
    return fail(ZS_EINVAL, "the segment count must be a power of two");
  uint32_t niv = 0;
  while ((1u << niv) < a.num_segments) ++niv;
  if (a.row_variables < 1)
    return fail(ZS_EINVAL, "eval_point must carry at least one row variable");
  if (a.row_variables + niv > ZS_MAX_ROUNDS) return fail(ZS_EINVAL, "too many variables");
  for (uint32_t s = 0; s < a.num_segments; ++s)
    if (a.row_counts[s] < 1 || a.row_counts[s] > (1u << a.row_variables))
      return fail(ZS_EINVAL, "row count " + std::to_string(a.row_counts[s]) +
                                 " outside [1, 2^" + std::to_string(a.row_variables) +
                                 "]: the row-eq lookup would run out of bounds");
  if (!known_kind(transcript->kind)) return fail(ZS_EINVAL, "unknown transcript kind");
  const uint32_t tf = transcript->field;
  if (!(tf == a.field || tf == base_of(a.field)))
    return fail(ZS_EINVAL, "challenge field does not tile the transcript field");
  if (transcript->item_bytes != field_bytes(tf))
    return fail(ZS_EINVAL, "transcript item size mismatch");
  int rc;
  if ((rc = set_device(ctx))) return rc;
  return prove_jagged_layer(ctx, a, transcript, out);
}

int zs_prove_batch(zs_ctx* ctx, size_t count, const zs_prove_args* args,
                   zs_transcript* transcripts, zs_proof* outs) {
  if (!ctx || (count && (!args || !transcripts || !outs))) return fail(ZS_EINVAL, "null argument");
  int rc;
  if ((rc = set_device(ctx))) return rc;
  // One group per kind (field pair, factor count, mode); every step launches
  // the round of each group's members that are still running.
  std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, bool>, std::unique_ptr<BatchGroup>>
      groups;
  for (size_t i = 0; i < count; ++i) {
    uint32_t npts = 0;
    if ((rc = validate_args(args[i], &transcripts[i], &npts))) return rc;
    if (args[i].refold_rounds)
      return fail(ZS_EINVAL, "batched proving folds every round (refold_rounds must be 0)");
    if (args[i].mode == ZS_MODE_LOGUP)
      return fail(ZS_EINVAL, "LogUp stages are proved one at a time (zs_prove)");
    const bool eq = args[i].mode == ZS_MODE_EQ;
    auto key = std::make_tuple(args[i].table_field, args[i].chal_field, args[i].num_tables, npts, eq);
    auto& g = groups[key];
    if (!g) {
      g.reset(make_batch_group(args[i].table_field, args[i].chal_field, args[i].num_tables, npts, eq));
      if (!g) return fail(ZS_EINVAL, "unsupported field / degree combination");
    }
    g->add(i, args[i], transcripts[i]);
  }
  uint32_t steps = 0;
  for (auto& kv : groups) {
    if ((rc = kv.second->setup(ctx, ctx->stream))) return rc;
    steps = std::max(steps, kv.second->max_rounds());
  }
  for (uint32_t st = 0; st < steps; ++st)
    for (auto& kv : groups)
      if ((rc = kv.second->enqueue_step(st, ctx->stream))) return rc;
  for (auto& kv : groups)
    if ((rc = kv.second->download(ctx->stream))) return rc;
  ZS_HIP(hipStreamSynchronize(ctx->stream));
  for (auto& kv : groups)
    if ((rc = kv.second->collect(outs, transcripts))) return rc;
  return ZS_OK;
}

// ------------------------------------------------------------ per round
int zs_state_create(zs_ctx* ctx, const zs_prove_args* args, zs_state** out) {
  if (!ctx || !args || !out) return fail(ZS_EINVAL, "null argument");
  uint32_t npts = 0;
  int rc = validate_args(*args, nullptr, &npts);
  if (rc) return rc;
  if ((rc = set_device(ctx))) return rc;
  auto st = std::make_unique<zs_state>();
  st->ctx = ctx;
  st->inst.reset(args->mode == ZS_MODE_LOGUP
                     ? make_logup_instance(args->chal_field)
                     : make_instance(args->table_field, args->chal_field, args->num_tables, npts,
                                     args->mode == ZS_MODE_EQ));
  if (!st->inst) return fail(ZS_EINVAL, "unsupported field / degree combination");
  zs_transcript blank;
  memset(&blank, 0, sizeof(blank));
  blank.item_bytes = field_bytes(base_of(args->chal_field));
  if ((rc = st->inst->setup(ctx, *args, blank))) return rc;
  ZS_HIP(hipStreamSynchronize(ctx->stream));
  *out = st.release();
  return ZS_OK;
}

int zs_state_round(zs_state* st, zs_transcript* t, void* msg_out, void* chal_out,
                   void* claim_out) {
  if (!st || !t) return fail(ZS_EINVAL, "null argument");
  if (st->bound >= st->inst->rounds) return fail(ZS_EINVAL, "every round is already bound");
  if (!known_kind(t->kind)) return fail(ZS_EINVAL, "unknown transcript kind");
  if (t->item_bytes == 0 || st->inst->chal_bytes % t->item_bytes)
    return fail(ZS_EINVAL, "challenge field does not tile the transcript field");
  int rc;
  if ((rc = set_device(st->ctx))) return rc;
  hipStream_t s = st->ctx->stream;
  if ((rc = st->inst->upload_transcript(*t, s))) return rc;
  if ((rc = st->inst->enqueue(st->bound, st->bound + 1, s))) return rc;
  if ((rc = st->inst->download_round(st->bound, s))) return rc;
  ZS_HIP(hipStreamSynchronize(s));
  if ((rc = st->inst->collect_round(st->bound, msg_out, chal_out, claim_out, t))) return rc;
  st->bound += 1;
  return ZS_OK;
}

int zs_state_bound(const zs_state* st, uint32_t* bound) {
  if (!st || !bound) return fail(ZS_EINVAL, "null argument");
  *bound = st->bound;
  return ZS_OK;
}

int zs_state_read_tables(zs_state* st, void* out, size_t out_bytes) {
  if (!st || !out) return fail(ZS_EINVAL, "null argument");
  int rc;
  if ((rc = set_device(st->ctx))) return rc;
  return st->inst->read_tables(st->bound, out, out_bytes, st->ctx->stream);
}

void zs_state_destroy(zs_state* st) {
  if (!st) return;
  (void)hipSetDevice(st->ctx->device);
  (void)hipStreamSynchronize(st->ctx->stream);
  delete st;
}

int zs_field_op(zs_ctx* ctx, uint32_t field, uint32_t op, const void* a, const void* b,
                void* out, size_t count) {
  if (!ctx || (count && (!a || !b || !out))) return fail(ZS_EINVAL, "null argument");
  if (op != ZS_OP_ADD && op != ZS_OP_SUB && op != ZS_OP_MUL) return fail(ZS_EINVAL, "unknown op");
  if (count == 0) return ZS_OK;
  if (count > ((size_t)1 << 28)) return fail(ZS_EINVAL, "at most 2^28 operations per call");
  int rc;
  if ((rc = set_device(ctx))) return rc;
  switch (field) {
    case ZS_F_BABYBEAR: return field_op_typed<BB>(ctx, op, a, b, out, count);
    case ZS_F_BABYBEAR4: return field_op_typed<BB4>(ctx, op, a, b, out, count);
    case ZS_F_GOLDILOCKS: return field_op_typed<GL>(ctx, op, a, b, out, count);
  }
  return fail(ZS_EINVAL, "unknown field id");
}

int zs_hbm_bandwidth(zs_ctx* ctx, size_t bytes, uint32_t reps, double* read_gbs,
                     double* copy_gbs) {
  if (!ctx || !read_gbs || !copy_gbs || reps == 0) return fail(ZS_EINVAL, "bad argument");
  bytes &= ~(size_t)4095;
  if (bytes < ((size_t)1 << 20)) return fail(ZS_EINVAL, "measure at least 1 MiB");
  int rc;
  if ((rc = set_device(ctx))) return rc;
  int dev = 0, cus = 0;
  ZS_HIP(hipGetDevice(&dev));
  ZS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const uint32_t grid = (uint32_t)cus * 8;
  void *a = nullptr, *b = nullptr, *sink = nullptr;
  if ((rc = ctx->pool->alloc(bytes, &a))) return rc;
  if ((rc = ctx->pool->alloc(bytes, &b))) {
    ctx->pool->release(a);
    return rc;
  }
  if ((rc = ctx->pool->alloc((size_t)grid * 4, &sink))) {
    ctx->pool->release(a);
    ctx->pool->release(b);
    return rc;
  }
  hipStream_t s = ctx->stream;
  hipEvent_t e0 = ctx->event(), e1 = ctx->event();
  const size_t n = bytes / 16;
  double best_read = 0, best_copy = 0;
  hipError_t e = hipMemsetAsync(a, 0x5a, bytes, s);
  for (uint32_t r = 0; r < reps + 1 && e == hipSuccess; ++r) {  // the first run warms up
    (void)hipEventRecord(e0, s);
    stream_read_kernel<<<grid, 256, 0, s>>>((const uint4*)a, n, (uint32_t*)sink);
    (void)hipEventRecord(e1, s);
    e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && r > 0 && ms > 0) best_read = std::max(best_read, bytes / (ms * 1e6));
    (void)hipEventRecord(e0, s);
    stream_copy_kernel<<<grid, 256, 0, s>>>((const uint4*)a, (uint4*)b, n);
    (void)hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && r > 0 && ms > 0)
      best_copy = std::max(best_copy, 2.0 * bytes / (ms * 1e6));
  }
  ctx->spare_events.push_back(e0);
  ctx->spare_events.push_back(e1);
  ctx->pool->release(a);
  ctx->pool->release(b);
  ctx->pool->release(sink);
  if (e != hipSuccess) return hip_fail(e, "zs_hbm_bandwidth");
  *read_gbs = best_read;
  *copy_gbs = best_copy;
  return ZS_OK;
}

int zs_verify_round(uint32_t chal_field, zs_transcript* t, const void* claim, const void* msg,
                    uint32_t npts, uint32_t degree, void* chal_out, void* reduced_out, int* ok) {
  if (!t || !claim || !msg || !ok) return fail(ZS_EINVAL, "null argument");
  if (degree < 1) return fail(ZS_EINVAL, "degree must be >= 1");
  if (npts != degree + 1)
    return fail(ZS_EINVAL, "round message must have " + std::to_string(degree + 1) +
                               " evals, got " + std::to_string(npts));
  if (!known_kind(t->kind)) return fail(ZS_EINVAL, "unknown transcript kind");
  const uint32_t cb = field_bytes(chal_field);
  if (!cb) return fail(ZS_EINVAL, "unknown field id");
  if (!(t->field == chal_field || t->field == base_of(chal_field)) || t->item_bytes == 0 ||
      cb % t->item_bytes)
    return fail(ZS_EINVAL, "challenge field does not tile the transcript field");
  switch (chal_field) {
    case ZS_F_BABYBEAR: return verify_round_typed<BB>(t, claim, msg, npts, chal_out, reduced_out, ok);
    case ZS_F_BABYBEAR4: return verify_round_typed<BB4>(t, claim, msg, npts, chal_out, reduced_out, ok);
    case ZS_F_GOLDILOCKS: return verify_round_typed<GL>(t, claim, msg, npts, chal_out, reduced_out, ok);
  }
  return fail(ZS_EINVAL, "unknown field id");
}

}  // extern "C"
