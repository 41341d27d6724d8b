// Proof instances, batch groups and server kinds with BB tables and BB challenges.
#include "batch.h"
#include "instance.h"
#include "serve.h"

namespace zs {
Instance* make_instance_bb(uint32_t k, uint32_t npts, bool eq) {
  return make_typed<BB, BB>(k, npts, eq);
}
BatchGroup* make_batch_group_bb(uint32_t k, uint32_t npts, bool eq) {
  return make_typed_batch<BB, BB>(k, npts, eq);
}
ServeKind* make_serve_kind_bb(uint32_t k, uint32_t npts, bool eq) {
  return make_typed_serve<BB, BB>(k, npts, eq);
}
}  // namespace zs
