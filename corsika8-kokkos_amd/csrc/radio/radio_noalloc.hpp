// One-shot radio projection without device allocation (radio_noalloc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "c8mi/c8mi_types.h"

namespace c8mi::radio {

// Projects `count` compact tracks ([9][stride] FP64, device) on the observers
// of `config` (deterministic mode) into the caller's device waveforms,
// [3][total bins] int64 per enabled algorithm, which it clears first. Device
// copies of the observers, the propagation table and the counters live in the
// library's fixed scratch array (C8MI_RADIO_NOALLOC_MAX_* bound them). Returns
// after the work on `stream` has completed; throws OverflowError when the
// fixed-point range was exceeded (the waveforms are then invalid).
void project_fixed_noalloc(c8mi_radio_config const& config, double const* device_tracks,
                           std::uint64_t stride, std::uint64_t count, std::int64_t* device_coreas,
                           std::int64_t* device_zhs, c8mi_radio_counters* counters,
                           hipStream_t stream);

}  // namespace c8mi::radio
