// Opaque C-ABI handles shared between the EM and radio entry points (an EM
// backend can project its lepton steps onto a radio handle).
#pragma once

#include <hip/hip_runtime.h>

#include "c8mi/c8mi_radio.h"
#include "radio/radio_accumulator.hpp"

struct c8mi_radio {
  c8mi::radio::RadioAccumulator accumulator;
  c8mi_radio(c8mi_radio_config const& config, hipStream_t stream) : accumulator(config, stream) {}
};
