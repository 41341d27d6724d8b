// Proof instances and batch groups with BB tables and BB challenges.
#include "batch.h"
#include "instance.h"

namespace zs {
Instance* make_instance_bb(uint32_t k, uint32_t npts, bool eq) {
  return make_typed<BB, BB>(k, npts, eq);
}
BatchGroup* make_batch_group_bb(uint32_t k, uint32_t npts, bool eq) {
  return make_typed_batch<BB, BB>(k, npts, eq);
}
}  // namespace zs
