// Host runtime of the prover: error reporting, the per-device context (one
// HIP stream, a budgeted device-memory pool) and the type-erased proof
// instance the C ABI drives.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/zorch_amd.h"

namespace zs {

// Thread-local message for the last failing call.
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define ZS_HIP(call)                                    \
  do {                                                  \
    hipError_t zs_e_ = (call);                          \
    if (zs_e_ != hipSuccess) return hip_fail(zs_e_, #call); \
  } while (0)

// Budgeted device-memory pool. Blocks are cached by exact size and reused;
// `budget` bounds live + cached bytes (cached blocks are released first when
// a new allocation would exceed it). Thread-safe: a proof server's scheduler
// thread and the caller's thread may share one context's pool.
class DevicePool {
 public:
  explicit DevicePool(uint64_t budget) : budget_(budget) {}
  ~DevicePool();
  int alloc(size_t bytes, void** out);
  void release(void* p);
  void trim();
  uint64_t held() const {
    std::lock_guard<std::mutex> g(mu_);
    return live_ + cached_;
  }
  uint64_t live() const {
    std::lock_guard<std::mutex> g(mu_);
    return live_;
  }
  uint64_t budget() const { return budget_; }

 private:
  void trim_locked();
  mutable std::mutex mu_;
  uint64_t budget_;
  uint64_t live_ = 0, cached_ = 0;
  std::multimap<size_t, void*> free_;
  std::map<void*, size_t> sizes_;
};

// Pinned host staging blocks, cached by size (uploads/downloads of the small
// per-proof control and output regions go through one copy each).
// Thread-safe, as DevicePool.
class HostPool {
 public:
  ~HostPool();
  int alloc(size_t bytes, void** out);
  void release(void* p);

 private:
  std::mutex mu_;
  std::multimap<size_t, void*> free_;
  std::map<void*, size_t> sizes_;
};

constexpr uint32_t TUNE_MAX_GRID = 4096;  // equals the kernels' MAXG

struct TuneParams {
  uint32_t tail_pairs = 1024;   // rounds with at most this many pairs run in the tail kernel
  uint32_t max_grid = 512;      // blocks of a multi-block pass (<= MAXG)
  uint32_t pairs_per_thread = 1;  // target pairs per thread of a multi-block pass
  uint32_t chain_pairs = 1u << 18;  // rounds >= 2 with at most this many pairs (and more than
                                    // tail_pairs) run chained in one persistent launch; 0: off
  uint32_t chain_grid = 128;        // blocks of a chained launch (at most)
  uint32_t step_rounds = 0;         // batched / served: tail rounds per step (0: all)
  uint32_t lookahead = 1;           // zs_prove: rounds 0-2 from the input; look-ahead rounds
};

// Kernel timing records (enabled by zs_ctx_profile).
struct KernelRecord {
  uint32_t round, kind, grid;
  uint64_t bytes;
  hipEvent_t start, stop;
};
enum {
  KREC_PASS = 1, KREC_FINISH = 2, KREC_TAIL = 3, KREC_BATCH = 4, KREC_CHAIN = 5,
  KREC_PASS_LA = 6, KREC_CHAIN_LA = 7, KREC_TAIL_LA = 8,  // look-ahead variants
  KREC_FUSED01 = 9, KREC_FOLD2 = 10,                      // rounds 0-2 from the input
  KREC_REFOLD = 11, KREC_ENGINE = 12,                     // sqrt-space rounds; other engines
  KREC_LOGUP = 13, KREC_JAGGED = 14, KREC_JAGGED_BATCH = 15  // LogUp-GKR layer rounds
};

}  // namespace zs

struct zs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  zs::DevicePool* pool = nullptr;
  zs::HostPool* host_pool = nullptr;
  zs::TuneParams tune;
  bool profile = false;
  uint64_t* trace = nullptr;  // device phase-trace buffer (null: off)
  std::vector<zs::KernelRecord> records;
  std::vector<hipEvent_t> spare_events;
  hipEvent_t event() {
    if (!spare_events.empty()) {
      hipEvent_t e = spare_events.back();
      spare_events.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
};

namespace zs {

// One proof in flight (type-erased over field / factor count / mode).
class Instance {
 public:
  virtual ~Instance() = default;
  // Allocate, upload inputs and the transcript, build eq tables (on
  // `stream`, with launch shapes `tune` for every later launch).
  virtual int setup(zs_ctx* ctx, const zs_prove_args& a, const zs_transcript& t,
                    hipStream_t stream, const TuneParams& tune) = 0;
  int setup(zs_ctx* ctx, const zs_prove_args& a, const zs_transcript& t) {
    return setup(ctx, a, t, ctx->stream, ctx->tune);
  }
  // As setup, in caller-provided device memory (dev_cap bytes) and host
  // staging (host_cap bytes, any host memory): allocates nothing; ZS_EINVAL if
  // the stage does not fit.
  virtual int setup_in(zs_ctx* ctx, const zs_prove_args& a, const zs_transcript& t,
                       hipStream_t stream, const TuneParams& tune, void* dev_mem, size_t dev_cap,
                       void* host_mem, size_t host_cap) = 0;
  // Enqueue rounds [first, last) on `stream`.
  virtual int enqueue(uint32_t first, uint32_t last, hipStream_t stream) = 0;
  // Enqueue the copy of the outputs to pinned host staging (whole proof).
  virtual int download(hipStream_t stream) = 0;
  // After the stream has synchronised: copy staged outputs to the caller.
  virtual int collect(zs_proof* out, zs_transcript* t) = 0;
  // Per-round API: upload a transcript state before a round; read one round.
  virtual int upload_transcript(const zs_transcript& t, hipStream_t stream) = 0;
  virtual int download_round(uint32_t round, hipStream_t stream) = 0;
  virtual int collect_round(uint32_t round, void* msg, void* chal, void* claim,
                            zs_transcript* t) = 0;
  virtual int read_tables(uint32_t bound, void* out, size_t out_bytes, hipStream_t stream) = 0;
  virtual void free_buffers() = 0;
  uint32_t rounds = 0;
  uint32_t npts = 0;
  uint32_t mlen = 0;  // message elements on the wire
  uint32_t chal_bytes = 0;
  uint32_t k = 0;
  uint64_t device_bytes = 0;
};

// Many proofs of one kind (field pair, factor count, mode) proved together:
// each step launches one round of every member still running.
class BatchGroup {
 public:
  virtual ~BatchGroup() = default;
  // Register member `index` (its args and transcript are read in setup).
  virtual void add(size_t index, const zs_prove_args& a, const zs_transcript& t) = 0;
  // Allocate, build the step schedule, upload control blocks and inputs.
  virtual int setup(zs_ctx* ctx, hipStream_t stream) = 0;
  virtual uint32_t max_rounds() const = 0;
  // Launch step `step` (round `step` of every member still running).
  virtual int enqueue_step(uint32_t step, hipStream_t stream) = 0;
  virtual int download(hipStream_t stream) = 0;
  // After the stream has synchronised: outputs of every member, by index.
  virtual int collect(zs_proof* outs, zs_transcript* ts) = 0;
  virtual void free_buffers() = 0;
  uint64_t device_bytes = 0;
};

// One admitted request of the proof server (device state lives with its kind).
class ServeReq {
 public:
  virtual ~ServeReq() = default;
  uint32_t n = 0;
  uint32_t next = 0;       // next round to run
  bool finishing = false;  // completes with the step just launched
  uint64_t bytes = 0;      // device bytes held
};

// Outputs of a finished request, kept on the host until polled.
struct ServeResult {
  std::vector<uint8_t> msgs, point, value, final_tables;
  std::vector<uint8_t> tr_state;  // the transcript stream state (zs_transcript after its header)
};

// Requests of one kind (field pair, factor count, mode): admission, one
// batched launch per server step, download and release.
class ServeKind {
 public:
  virtual ~ServeKind() = default;
  virtual uint64_t bytes_needed(const zs_prove_args& a) const = 0;
  virtual int admit(zs_ctx* ctx, const zs_prove_args& a, const zs_transcript& t,
                    ServeReq** out) = 0;
  // Launch the next round of every request in `reqs` (one step); advances
  // their `next` and sets `finishing` on those that complete with it.
  virtual int step(zs_ctx* ctx, std::vector<ServeReq*>& reqs, hipStream_t stream) = 0;
  virtual int download(ServeReq* r, hipStream_t stream) = 0;
  virtual void collect(ServeReq* r, ServeResult* out) = 0;
  virtual void release(zs_ctx* ctx, ServeReq* r) = 0;
};

// LogUp stages (logup.hip): tables and challenges in `field`; nullptr for
// an unsupported field.
Instance* make_logup_instance(uint32_t field);
// The dense LogUp-GKR circuit on the device (logup.hip): one transition of
// a 5-row layer buffer, and the eq table of a point into a row.
int logup_transition(zs_ctx* ctx, uint32_t field, const void* layer, size_t width, void* next);
int logup_eq(zs_ctx* ctx, uint32_t field, const void* point, uint32_t n, void* out);
// One transition of the jagged circuit on the device (jagged.hip).
int jagged_transition(zs_ctx* ctx, uint32_t field, uint32_t nseg, const uint32_t* rows,
                      const uint32_t* out_rows, const void* planes, size_t width, void* next,
                      size_t out_width);
// The jagged LogUp-GKR layer sumcheck (jagged.hip); arguments validated by
// zs_prove_jagged_layer.
int prove_jagged_layer(zs_ctx* ctx, const zs_jagged_args& a, zs_transcript* t,
                       zs_jagged_proof* out);
// Factory: returns nullptr for an unsupported combination.
Instance* make_instance(uint32_t table_field, uint32_t chal_field, uint32_t k, uint32_t npts,
                        bool eq);
ServeKind* make_serve_kind(uint32_t table_field, uint32_t chal_field, uint32_t k, uint32_t npts,
                           bool eq);
BatchGroup* make_batch_group(uint32_t table_field, uint32_t chal_field, uint32_t k,
                             uint32_t npts, bool eq);
// Per field pair and factor count (one translation unit each, units.h).
#define ZS_UNIT_DECL(NAME, KK)                                            \
  Instance* make_instance_##NAME##_k##KK(uint32_t npts, bool eq);         \
  BatchGroup* make_batch_group_##NAME##_k##KK(uint32_t npts, bool eq);    \
  ServeKind* make_serve_kind_##NAME##_k##KK(uint32_t npts, bool eq);
#define ZS_UNIT_DECL4(NAME) \
  ZS_UNIT_DECL(NAME, 1) ZS_UNIT_DECL(NAME, 2) ZS_UNIT_DECL(NAME, 3) ZS_UNIT_DECL(NAME, 4)
ZS_UNIT_DECL4(bb)
ZS_UNIT_DECL4(bb_bb4)
ZS_UNIT_DECL4(bb4)
ZS_UNIT_DECL4(gl)
#undef ZS_UNIT_DECL4
#undef ZS_UNIT_DECL

// Message width of a wire form for a round polynomial of npts natural values
// (0 with an error set when the form does not apply). `product_two`: a
// two-factor product round (the compressed form's only case).
inline uint32_t wire_width(const zs_wire* w, uint32_t npts, bool product_two) {
  if (!w || w->kind == ZS_WIRE_VALUES) return npts;
  switch (w->kind) {
    case ZS_WIRE_DOMAIN: {
      const uint32_t m = w->num_nodes + (w->has_inf ? 1u : 0u);
      if (m == 0) {
        set_error("an evaluation domain needs at least one sample point");
        return 0;
      }
      if (m > npts) {
        set_error("the evaluation domain has " + std::to_string(m) +
                  " samples; a round polynomial of degree " + std::to_string(npts - 1) +
                  " is sent with at most " + std::to_string(npts));
        return 0;
      }
      if (w->has_inf && w->inf_pos > w->num_nodes) {
        set_error("inf_pos must be at most the number of finite nodes");
        return 0;
      }
      if (w->num_nodes && !w->nodes) {
        set_error("null evaluation-domain nodes");
        return 0;
      }
      return m;
    }
    case ZS_WIRE_COEFFS:
      return npts;
    case ZS_WIRE_COMPRESSED:
      if (!product_two) {
        set_error("compressed product round takes exactly 2 factors");
        return 0;
      }
      return 2;
  }
  set_error("unknown wire form");
  return 0;
}

// Device address of the no-allocation scratch array on `device` (current),
// looked up once per device (noalloc.hip).
int noalloc_scratch_init(int device);

uint32_t field_bytes(uint32_t field);
uint32_t base_of(uint32_t field);  // base prime field of a field id
int set_device(zs_ctx* ctx);
// One host Fiat-Shamir hop (observe under slice framing, sample `limbs`
// words) on either transcript kind; false if a BLAKE3 stream outgrows its
// subtree stack.
bool transcript_hop(zs_transcript* t, const uint8_t* msg, size_t len, size_t item, size_t limbs,
                    uint8_t* out);

// Structural checks of a request: the natural value count npts and the
// message width mlen on its wire form.
int validate_args(const zs_prove_args& a, const zs_transcript* t, uint32_t* npts,
                  uint32_t* mlen = nullptr);

}  // namespace zs
