// Proof instances with BB tables and BB4 challenges.
#include "instance.h"

namespace zs {
Instance* make_instance_bb_bb4(uint32_t k, uint32_t npts, bool eq) {
  return make_typed<BB, BB4>(k, npts, eq);
}
}  // namespace zs
