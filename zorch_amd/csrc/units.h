// One instantiation unit: the proof instance, batch group and server kind of
// one field pair and factor count, with and without the eq factor.
#pragma once
#include "batch.h"
#include "instance.h"
#include "serve.h"

#define ZS_UNIT(NAME, FT, FC, KK)                                              \
  namespace zs {                                                                \
  Instance* make_instance_##NAME##_k##KK(uint32_t npts, bool eq) {              \
    return make_typed_k<FT, FC, KK>(npts, eq);                                  \
  }                                                                             \
  BatchGroup* make_batch_group_##NAME##_k##KK(uint32_t npts, bool eq) {         \
    return make_typed_batch_k<FT, FC, KK>(npts, eq);                            \
  }                                                                             \
  ServeKind* make_serve_kind_##NAME##_k##KK(uint32_t npts, bool eq) {           \
    return make_typed_serve_k<FT, FC, KK>(npts, eq);                            \
  }                                                                             \
  }
