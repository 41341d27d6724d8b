// Proof instances, batch groups and server kinds with BB4 tables and BB4 challenges.
#include "batch.h"
#include "instance.h"
#include "serve.h"

namespace zs {
Instance* make_instance_bb4(uint32_t k, uint32_t npts, bool eq) {
  return make_typed<BB4, BB4>(k, npts, eq);
}
BatchGroup* make_batch_group_bb4(uint32_t k, uint32_t npts, bool eq) {
  return make_typed_batch<BB4, BB4>(k, npts, eq);
}
ServeKind* make_serve_kind_bb4(uint32_t k, uint32_t npts, bool eq) {
  return make_typed_serve<BB4, BB4>(k, npts, eq);
}
}  // namespace zs
