// Per-shower radio of a pool of concurrent showers: every pool slot keeps its
// own observers, int64 fixed-point waveforms and counters resident on the
// device, and one projection launch per batch of tracks covers every slot's
// tracks on that slot's observers only.
//
// The reference gives each shower its own RadioAccumulator (one per backend
// instance, BackendConfig::radio_, corsika/kokkos/em/execution/BackendConfig.hpp:66)
// fed per lepton front (LeptonCascade.inl:81-82); here the showers of the pool
// share the fronts, so their accumulators share the launch: a batch's tracks
// come grouped in segments of one slot each (SegmentDev), and the projection
// kernel maps each workgroup to (segment, track chunk, observer of the
// segment's slot, window tile).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <vector>

#include "c8mi/c8mi_types.h"
#include "../runtime.hpp"
#include "radio_device.hpp"

namespace c8mi::radio {

class SlotRadio {
 public:
  // config: algorithms, refinement, exactness, fixed-point limit and
  // propagation (its observer lists are ignored); deterministic (int64) only.
  SlotRadio(c8mi_radio_config const& config, std::uint32_t slots, std::uint32_t observers_per_slot,
            std::uint32_t bins_per_observer, hipStream_t stream);
  ~SlotRadio();
  SlotRadio(SlotRadio const&) = delete;
  SlotRadio& operator=(SlotRadio const&) = delete;

  hipStream_t stream() const noexcept { return stream_; }
  std::uint32_t observers_per_slot() const noexcept { return observers_per_slot_; }
  // Tracks per workgroup for a batch of `count` tracks: about kTargetItems
  // items over all observers, at least kMinimumChunk tracks each (a
  // workgroup clears and flushes a whole window; fewer tracks would not pay
  // for it).
  std::uint64_t chunk_for(std::uint64_t count) const noexcept;
  std::uint32_t bin_tiles() const noexcept { return bin_tiles_; }
  std::uint32_t const* observer_counts() const noexcept { return observer_counts_.data(); }

  // Replaces a slot's observers (count <= observers_per_slot, bins <=
  // bins_per_observer; none: the slot's tracks are not projected) and clears
  // its waveforms and counters, after the projections already queued.
  void set_slot(std::uint32_t slot, c8mi_radio_observer const* observers, std::uint64_t count);
  // The slot's raw fixed-point sums, [3][sum of its observers' bins] per
  // algorithm (either may be null), and counters; waits for the queued
  // projections.
  void download(std::uint32_t slot, std::int64_t* coreas, std::int64_t* zhs, c8mi_radio_counters* counters);
  // Sum of the bins of a slot's observers.
  std::uint64_t slot_bins(std::uint32_t slot) const;

  // Projects tracks [9][stride] of several slots: `segments` (device, at most
  // max_segments, *segment_count of them) give each slot's track ranges and
  // item offsets (chunks x observers x tiles), *item_count their total.
  // Asynchronous on stream().
  // Throws what project() of `count` tracks in chunks of `chunk` over at most
  // `max_segments` segments would throw, without launching anything.
  void check_project(std::uint64_t count, std::uint64_t chunk, std::uint64_t max_segments) const;
  void project(double const* tracks, std::uint64_t stride, std::uint64_t count, std::uint64_t chunk,
               SegmentDev const* segments, std::uint32_t const* segment_count,
               unsigned long long const* item_count, std::uint32_t max_segments);

  void enable_kernel_timing(bool enable);
  // Projection-kernel time (ms) and launches since timing was enabled.
  void kernel_time_ms(double& ms, std::uint64_t& launches);

 private:
  std::uint64_t grid_for(std::uint64_t count, std::uint64_t chunk, std::uint64_t max_segments) const;
  c8mi_radio_config config_{};
  std::uint32_t slots_ = 0;
  std::uint32_t observers_per_slot_ = 0;
  std::uint32_t bins_per_observer_ = 0;
  std::uint64_t total_bins_ = 0;
  std::uint32_t tile_bins_ = 0;
  std::uint32_t bin_tiles_ = 1;
  bool fuse_ = false;
  bool owns_stream_ = false;
  hipStream_t stream_ = nullptr;
  PropagationDev propagation_{};
  DeviceBuffer<double> refractivity_;
  DeviceBuffer<double> integrated_;
  DeviceBuffer<ObserverDev> coreas_observers_;
  DeviceBuffer<ObserverDev> zhs_observers_;
  DeviceBuffer<std::uint32_t> observer_counts_;
  DeviceBuffer<long long> coreas_fixed_;
  DeviceBuffer<long long> zhs_fixed_;
  DeviceBuffer<CountersDev> counters_;
  std::vector<std::vector<c8mi_radio_observer>> slot_observers_;
  std::vector<std::uint32_t> counts_host_;
  std::vector<ObserverDev> staging_;
  std::vector<long long> download_staging_;
  bool timing_ = false;
  std::vector<std::array<hipEvent_t, 2>> pending_events_;
  double ms_ = 0.;
  std::uint64_t launches_ = 0;
};

}  // namespace c8mi::radio
