// Device-side types shared by the radio kernels and their host launcher.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace c8mi::radio {

// One observer as the kernels see it (RadioObserverView,
// corsika/kokkos/radio/RadioProjectionData.hpp:39-49, with the waveform offset
// shared by the x/y/z blocks).
struct ObserverDev {
  double position[3];
  double start_time;
  double duration;
  double sample_rate;
  double fixed_scale;  // 2^62 * (ZHS ? sample_rate : 1) / field limit
  std::uint64_t number_of_bins;
  std::uint64_t waveform_offset;
};

// Straight-ray propagation table (RadioPropagationView,
// RadioProjectionData.hpp:50-64) with device pointers.
struct PropagationDev {
  double homogeneous_index;
  double minimum_height;
  double maximum_height;
  double step;
  double inverse_step;
  double slope_n_lower;
  double slope_i_lower;
  double slope_n_upper;
  double slope_i_upper;
  double const* refractivity;
  double const* integrated;
  std::uint64_t table_size;
};

// Global counters (RadioCounters' projection subset) plus the conservative
// wrap detector used by the LDS difference-array accumulation.
struct CountersDev {
  unsigned long long coreas_contributions;
  unsigned long long zhs_contributions;
  unsigned long long zhs_subtracks;
  unsigned long long valid_tracks;
  unsigned long long fixed_point_overflows;
  unsigned long long pad[3];
};

// Tracks [begin, end) of one pool slot in a batch of several showers' tracks;
// its items (chunks x the slot's observers x window tiles) start at item_begin.
struct SegmentDev {
  std::uint64_t begin;
  std::uint64_t end;
  std::uint64_t item_begin;
  std::uint32_t slot;
  std::uint32_t pad;
};

struct ProjectionLaunch {
  double const* tracks;  // [9][stride] compact tracks
  std::uint64_t stride;
  std::uint64_t track_count;
  ObserverDev const* coreas_observers;  // may equal zhs_observers (fused)
  ObserverDev const* zhs_observers;
  std::uint32_t observer_count;
  std::uint32_t zhs_refinement;
  PropagationDev propagation;
  void* coreas_waves;  // int64 (deterministic) or double, 3 * total_bins
  void* zhs_waves;
  std::uint64_t coreas_total_bins;
  std::uint64_t zhs_total_bins;
  CountersDev* counters;
  std::uint64_t chunk_size;  // tracks per workgroup
  std::uint32_t chunks;
  std::uint32_t bin_tiles;   // window tiles per observer
  std::uint32_t tile_bins;   // bins held in LDS per mode
  std::uint32_t count_valid; // tally valid tracks in this launch (once per batch)
  std::uint32_t pad;
  // Pooled showers (segments non-null): workgroup b < *item_count is an item of
  // the segment holding it; the segment's slot selects the observers
  // [slot * slot_observers, + slot_observer_counts[slot]) and counters[slot].
  SegmentDev const* segments;
  std::uint32_t const* segment_count;
  unsigned long long const* item_count;
  std::uint32_t const* slot_observer_counts;
  std::uint32_t slot_observers;
  std::uint32_t pad2;
};

}  // namespace c8mi::radio
