// c8mi::radio::RadioAccumulator — MI355X replacement for
// corsika::kokkos::radio::RadioAccumulator<Exec>
// (corsika/kokkos/radio/RadioAccumulator.hpp:35-237).
//
// Same lifecycle: construct (initialize) -> reset -> any number of
// enqueue_host_lepton_tracks / accumulate_* calls -> download or download_fixed.
// All work is ordered on one HIP stream (the reference's single execution
// instance, RadioAccumulator.hpp:28-34); host enqueue copies the records before
// returning.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "c8mi/c8mi_types.h"
#include "../runtime.hpp"
#include "radio_device.hpp"

namespace c8mi::radio {

struct RadioStatistics {
  std::uint64_t host_tracks_enqueued = 0;
  std::uint64_t host_track_batches = 0;
  std::uint64_t projection_launches = 0;
  std::uint64_t track_observer_pairs = 0;
  std::uint64_t fused_track_observer_pairs = 0;
  std::uint64_t host_to_device_bytes = 0;
  std::uint64_t device_to_host_bytes = 0;
};

class RadioAccumulator {
 public:
  // HostTrackBatchCapacity (RadioAccumulator.hpp:49).
  static constexpr std::size_t kHostTrackBatchCapacity = 4096;

  // Validates and uploads `config` (RadioAccumulator::initialize and validate,
  // RadioAccumulator.inl:140-200, 934-975). `stream` 0 creates an owned stream.
  RadioAccumulator(c8mi_radio_config const& config, hipStream_t stream);
  ~RadioAccumulator();
  RadioAccumulator(RadioAccumulator const&) = delete;
  RadioAccumulator& operator=(RadioAccumulator const&) = delete;

  void reset();
  void enqueue_host_lepton_tracks(c8mi_lepton_transport_record const* records,
                                  std::size_t count);
  void flush_host_lepton_tracks();
  // Device-resident LeptonTransportRecords (accumulateLeptonTracks).
  void accumulate_lepton_records(c8mi_lepton_transport_record const* device_records,
                                 std::size_t count);
  // Device-resident compact tracks, [9][stride] FP64 (c8mi_types.h).
  void accumulate_tracks(double const* device_tracks, std::uint64_t stride,
                         std::uint64_t count);
  // Throws what accumulate_tracks of `count` valid tracks would throw, without
  // launching anything (callers check before work that cannot be undone).
  void check_accumulate(std::uint64_t count) const;

  // Decoded waveforms, [3][total_bins] per algorithm (x, y, z blocks). Either
  // output may be null. Throws OverflowError when the fixed-point range was
  // exceeded (RadioAccumulator::download, RadioAccumulator.inl:443-470).
  void download(double* coreas, double* zhs, c8mi_radio_counters* counters);
  // Raw fixed-point sums (downloadFixed); deterministic mode only.
  void download_fixed(std::int64_t* coreas, std::int64_t* zhs, c8mi_radio_counters* counters);

  // Launch shape: tracks per workgroup (0 = automatic), LDS window tile in bins
  // (0 = automatic), target workgroups per launch, threads per workgroup
  // (256/512/1024) and whether fusable CoREAS+ZHS work is split into two launches.
  void set_tuning(std::uint64_t chunk_tracks, std::uint32_t tile_bins,
                  std::uint32_t target_workgroups, std::uint32_t block_threads, bool split);
  void enable_kernel_timing(bool enable);
  // Projection-kernel durations measured with HIP events on this
  // accumulator's stream since timing was enabled, per launch kind:
  // [0] fused CoREAS+ZHS, [1] CoREAS only, [2] ZHS only.
  void kernel_time_ms(double ms[3], std::uint64_t launches[3]);

  std::uint64_t coreas_total_bins() const noexcept { return coreas_bins_; }
  std::uint64_t zhs_total_bins() const noexcept { return zhs_bins_; }
  bool fused() const noexcept { return fuse_; }
  hipStream_t stream() const noexcept { return stream_; }
  RadioStatistics const& statistics() const noexcept { return statistics_; }

 private:
  void project(double const* tracks, std::uint64_t stride, std::uint64_t count);
  void ensure_track_capacity(std::uint64_t count);
  ProjectionLaunch base_launch(bool coreas, bool zhs, std::uint32_t observers,
                               std::uint64_t max_bins, std::uint64_t count) const;
  void copy_counters(c8mi_radio_counters* out);

  c8mi_radio_config config_{};
  bool fuse_ = false;
  bool owns_stream_ = false;
  hipStream_t stream_ = nullptr;
  bool downloaded_ = false;
  bool sealed_ = false;

  std::vector<ObserverDev> coreas_host_;
  std::vector<ObserverDev> zhs_host_;
  std::vector<c8mi_radio_observer> coreas_snapshots_;
  std::vector<c8mi_radio_observer> zhs_snapshots_;
  std::uint64_t coreas_bins_ = 0;
  std::uint64_t zhs_bins_ = 0;
  std::uint64_t coreas_max_bins_ = 0;
  std::uint64_t zhs_max_bins_ = 0;

  DeviceBuffer<ObserverDev> coreas_observers_;
  DeviceBuffer<ObserverDev> zhs_observers_;
  DeviceBuffer<double> refractivity_;
  DeviceBuffer<double> integrated_;
  PropagationDev propagation_{};
  DeviceBuffer<long long> coreas_fixed_;
  DeviceBuffer<long long> zhs_fixed_;
  DeviceBuffer<double> coreas_floating_;
  DeviceBuffer<double> zhs_floating_;
  DeviceBuffer<CountersDev> counters_;
  DeviceBuffer<double> tracks_;
  std::uint64_t track_capacity_ = 0;
  DeviceBuffer<c8mi_lepton_transport_record> host_batch_device_;
  PinnedBuffer<c8mi_lepton_transport_record> host_batch_;
  std::size_t host_count_ = 0;

  std::uint64_t chunk_tracks_ = 0;     // 0: automatic
  std::uint32_t tile_bins_limit_ = 0;  // 0: automatic
  // Defaults measured on MI355X for 2^20 tracks x 160 antennas x 2048 bins:
  // separate CoREAS and ZHS launches (48 KiB LDS each, three workgroups per
  // CU), 512 threads, about 8192 workgroups per launch.
  std::uint32_t target_workgroups_ = 8192;
  std::uint32_t block_threads_ = 512;
  bool split_ = true;

  bool timing_ = false;
  struct TimedLaunch {
    hipEvent_t start, stop;
    int kind;
  };
  std::vector<TimedLaunch> pending_events_;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> free_events_;
  double timed_ms_[3] = {0., 0., 0.};
  std::uint64_t timed_launches_[3] = {0, 0, 0};

  RadioStatistics statistics_{};
};

}  // namespace c8mi::radio
