// Proof instances and batch groups with GL tables and GL challenges.
#include "batch.h"
#include "instance.h"

namespace zs {
Instance* make_instance_gl(uint32_t k, uint32_t npts, bool eq) {
  return make_typed<GL, GL>(k, npts, eq);
}
BatchGroup* make_batch_group_gl(uint32_t k, uint32_t npts, bool eq) {
  return make_typed_batch<GL, GL>(k, npts, eq);
}
}  // namespace zs
