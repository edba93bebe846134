// Validation and device forms of radio configurations, shared by the radio
// accumulator and the per-slot radio of the EM pool.
#pragma once

#include <cstdint>
#include <vector>

#include "c8mi/c8mi_types.h"
#include "radio_device.hpp"

namespace c8mi::radio {

inline constexpr double kHeadroom = 4611686018427387904.;  // 2^62, RadioProjectionData.hpp:34

// RadioAccumulator::validate (corsika/detail/kokkos/radio/RadioAccumulator.inl:934-975).
void validate_radio_config(c8mi_radio_config const& config);
// Its observer checks alone, against a validated propagation snapshot.
void validate_observers(c8mi_radio_observer const* observers, std::uint64_t count,
                        c8mi_radio_propagation const& p);
// RadioAccumulator::makeObservers (RadioAccumulator.inl:591-617): device
// observers with consecutive waveform offsets from 0.
std::vector<ObserverDev> make_observers(c8mi_radio_observer const* input, std::uint64_t count, double limit,
                                        bool zhs_potential, std::uint64_t& total_bins, std::uint64_t& max_bins);
// observersCanShareTrackKinematics (RadioAccumulator.inl:619-632).
bool can_fuse(std::vector<ObserverDev> const& a, std::vector<ObserverDev> const& b);
// The propagation snapshot with device table pointers.
PropagationDev propagation_view(c8mi_radio_propagation const& p, double const* refractivity,
                                double const* integrated);

}  // namespace c8mi::radio
