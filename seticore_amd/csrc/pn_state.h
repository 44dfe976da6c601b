// Host-side state of a block proposal network handle (shared by the C-ABI
// sources; not installed).
#pragma once

#include <cstddef>

#include "pn_internal.h"

struct sc_pn {
  int max_batch = 0;
  int device = 0;
  sc::pn::Weights w;
  sc::pn::Acts a;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  int step = 0;
  unsigned short* z = nullptr;  // max_batch tiles, staging for the gated search
  float* logits = nullptr;
  int* argmax = nullptr;
  float* dz = nullptr;
  float* gpart = nullptr;  // max_batch * kNumParams
  float* grad = nullptr;
  float* loss = nullptr;
  size_t device_bytes = 0;
};
