// Proof instances with GL tables and GL challenges.
#include "instance.h"

namespace zs {
Instance* make_instance_gl(uint32_t k, uint32_t npts, bool eq) {
  return make_typed<GL, GL>(k, npts, eq);
}
}  // namespace zs
