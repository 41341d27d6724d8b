// zs_prove_noalloc: a whole stage proved in a library-owned per-device scratch
// array, for callers that cannot allocate (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc). No counterpart in the
// reference.
//
// `scratch` is a file-scope device array; each device that loads the library
// holds its own copy until the process ends. Its name in host code is only a
// host-side stand-in, so its device address is looked up once per device with
// hipGetSymbolAddress (at context creation, with that device current) and only
// that address is handed to kernels and copies. The entry point itself calls
// no hipMalloc, hipFree or other device or pinned-host allocation: the proof's
// device block (control block, messages, slot rows, eq families, folded
// tables, staged input) lives in the scratch array, and its host mirror in a
// per-thread static buffer. Every call on a device shares that device's array
// and leaves its contents behind: the upload of the control block and the
// zeroed slot rows starts each call afresh. Calls that overlap on one device
// race on the array.
#include <string.h>

#include <memory>
#include <mutex>
#include <string>

#include "../../include/zorch_amd.h"
#include "runtime.h"

namespace zs {

static __device__ __align__(256) uint8_t scratch[ZS_NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES];

// Host mirror of a proof's io region (control block, outputs, slot rows, eq
// point): per thread, so calls from different threads on different devices
// do not share it.
constexpr size_t NOALLOC_HOST_BYTES = 64 * 1024;
alignas(256) static thread_local uint8_t t_host_io[NOALLOC_HOST_BYTES];

// The scratch array's device address on `device`, which must be current;
// looked up once per device.
int noalloc_scratch_init(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES)
    return fail(ZS_EINVAL, "device index beyond the no-allocation scratch table");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_ptr[device]) return ZS_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  if (!ptr) return fail(ZS_EHIP, "hipGetSymbolAddress(scratch) returned null");
  g_scratch_ptr[device] = ptr;
  return ZS_OK;
}

static void* scratch_ptr(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  return g_scratch_ptr[device];
}

}  // namespace zs

using namespace zs;

extern "C" int zs_prove_noalloc(zs_ctx* ctx, const zs_prove_args* args,
                                zs_transcript* transcript, zs_proof* out) {
  if (!ctx || !args || !transcript || !out) return fail(ZS_EINVAL, "null argument");
  uint32_t npts = 0;
  int rc = validate_args(*args, transcript, &npts);
  if (rc) return rc;
  if (args->refold_rounds)
    return fail(ZS_EINVAL, "the no-allocation entry point folds every round (refold_rounds must be 0)");
  if (args->mode == ZS_MODE_LOGUP)
    return fail(ZS_EINVAL, "the no-allocation entry point proves product summands");
  if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES)
    return fail(ZS_EINVAL, "device index beyond the no-allocation scratch table");
  ZS_HIP(hipSetDevice(ctx->device));
  void* dev = scratch_ptr(ctx->device);
  if (!dev) {
    // The context was created before the lookup could run (it runs in
    // zs_ctx_create); resolve it now, with this device current.
    if ((rc = noalloc_scratch_init(ctx->device))) return rc;
    dev = scratch_ptr(ctx->device);
  }
  std::unique_ptr<Instance> inst(make_instance(args->table_field, args->chal_field,
                                               args->num_tables, npts,
                                               args->mode == ZS_MODE_EQ));
  if (!inst) return fail(ZS_EINVAL, "unsupported field / degree combination");
  if ((rc = inst->setup_in(ctx, *args, *transcript, ctx->stream, ctx->tune, dev,
                           ZS_NOALLOC_SCRATCH_BYTES, t_host_io, NOALLOC_HOST_BYTES)))
    return rc;
  if ((rc = inst->enqueue(0, args->rounds, ctx->stream))) return rc;
  if ((rc = inst->download(ctx->stream))) return rc;
  ZS_HIP(hipStreamSynchronize(ctx->stream));
  return inst->collect(out, transcript);
}
