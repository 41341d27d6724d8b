// Proof instances with BB tables and BB challenges.
#include "instance.h"

namespace zs {
Instance* make_instance_bb(uint32_t k, uint32_t npts, bool eq) {
  return make_typed<BB, BB>(k, npts, eq);
}
}  // namespace zs
