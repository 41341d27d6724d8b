// Proof instances, batch groups and server kinds with GL tables and GL challenges.
#include "batch.h"
#include "instance.h"
#include "serve.h"

namespace zs {
Instance* make_instance_gl(uint32_t k, uint32_t npts, bool eq) {
  return make_typed<GL, GL>(k, npts, eq);
}
BatchGroup* make_batch_group_gl(uint32_t k, uint32_t npts, bool eq) {
  return make_typed_batch<GL, GL>(k, npts, eq);
}
ServeKind* make_serve_kind_gl(uint32_t k, uint32_t npts, bool eq) {
  return make_typed_serve<GL, GL>(k, npts, eq);
}
}  // namespace zs
