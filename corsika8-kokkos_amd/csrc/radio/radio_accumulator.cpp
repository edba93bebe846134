// Host side of the MI355X radio accumulator. See radio_accumulator.hpp.
#include "radio_accumulator.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>

#include "radio_launch.hpp"
#include "radio_setup.hpp"

namespace c8mi::radio {
namespace {

template <class T>
void upload(DeviceBuffer<T>& device, T const* host, std::size_t count, hipStream_t stream) {
  device.allocate(count);
  if (count)
    C8MI_HIP(hipMemcpyAsync(device.data(), host, count * sizeof(T), hipMemcpyHostToDevice,
                            stream));
}

}  // namespace

RadioAccumulator::RadioAccumulator(c8mi_radio_config const& config, hipStream_t stream)
    : config_(config) {
  validate_radio_config(config);
  if (stream == nullptr) {
    C8MI_HIP(hipStreamCreate(&stream_));
    owns_stream_ = true;
  } else {
    stream_ = stream;
  }
  try {
    double const limit = config.fixed_point_field_limit_v_per_m;
    if (config.coreas_enabled) {
      coreas_snapshots_.assign(config.coreas_observers,
                               config.coreas_observers + config.coreas_observer_count);
      coreas_host_ = make_observers(config.coreas_observers, config.coreas_observer_count,
                                    limit, false, coreas_bins_, coreas_max_bins_);
    }
    if (config.zhs_enabled) {
      zhs_snapshots_.assign(config.zhs_observers,
                            config.zhs_observers + config.zhs_observer_count);
      zhs_host_ = make_observers(config.zhs_observers, config.zhs_observer_count, limit, true,
                                 zhs_bins_, zhs_max_bins_);
    }
    // The config keeps borrowed pointers only for validation; the snapshots
    // above own the observer data from here on.
    config_.coreas_observers = nullptr;
    config_.zhs_observers = nullptr;
    upload(coreas_observers_, coreas_host_.data(), coreas_host_.size(), stream_);
    upload(zhs_observers_, zhs_host_.data(), zhs_host_.size(), stream_);
    auto const& p = config.propagation;
    if (!(p.homogeneous_refractive_index > 0.)) {
      upload(refractivity_, p.refractivity, p.table_size, stream_);
      upload(integrated_, p.integrated_refractivity, p.table_size, stream_);
    }
    propagation_ = propagation_view(p, refractivity_.data(), integrated_.data());
    config_.propagation.refractivity = nullptr;
    config_.propagation.integrated_refractivity = nullptr;
    if (config.deterministic) {
      coreas_fixed_.allocate(3 * coreas_bins_);
      zhs_fixed_.allocate(3 * zhs_bins_);
    } else {
      coreas_floating_.allocate(3 * coreas_bins_);
      zhs_floating_.allocate(3 * zhs_bins_);
    }
    counters_.allocate(1);
    host_batch_ = PinnedBuffer<c8mi_lepton_transport_record>(kHostTrackBatchCapacity);
    host_batch_device_.allocate(kHostTrackBatchCapacity);
    fuse_ = config.coreas_enabled && config.zhs_enabled && can_fuse(coreas_host_, zhs_host_);
    C8MI_HIP(hipStreamSynchronize(stream_));
    reset();
  } catch (...) {
    if (owns_stream_) (void)hipStreamDestroy(stream_);
    throw;
  }
}

RadioAccumulator::~RadioAccumulator() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& e : pending_events_) {
    (void)hipEventDestroy(e.start);
    (void)hipEventDestroy(e.stop);
  }
  for (auto& e : free_events_) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  if (owns_stream_) (void)hipStreamDestroy(stream_);
}

void RadioAccumulator::reset() {
  if (host_count_ != 0)
    throw std::logic_error("cannot reset radio with uncommitted host tracks");
  if (coreas_fixed_.size())
    C8MI_HIP(hipMemsetAsync(coreas_fixed_.data(), 0, coreas_fixed_.bytes(), stream_));
  if (zhs_fixed_.size())
    C8MI_HIP(hipMemsetAsync(zhs_fixed_.data(), 0, zhs_fixed_.bytes(), stream_));
  if (coreas_floating_.size())
    C8MI_HIP(hipMemsetAsync(coreas_floating_.data(), 0, coreas_floating_.bytes(), stream_));
  if (zhs_floating_.size())
    C8MI_HIP(hipMemsetAsync(zhs_floating_.data(), 0, zhs_floating_.bytes(), stream_));
  C8MI_HIP(hipMemsetAsync(counters_.data(), 0, counters_.bytes(), stream_));
  C8MI_HIP(hipStreamSynchronize(stream_));
  statistics_ = {};
  downloaded_ = false;
  sealed_ = false;
}

void RadioAccumulator::enqueue_host_lepton_tracks(c8mi_lepton_transport_record const* records,
                                                  std::size_t count) {
  if (downloaded_ || sealed_)
    throw std::logic_error("radio already downloaded; reset before host tracks");
  if (count != 0 && records == nullptr) throw std::invalid_argument("null host radio track input");
  while (count != 0) {
    std::size_t const copied = std::min(count, kHostTrackBatchCapacity - host_count_);
    std::memcpy(host_batch_.data() + host_count_, records,
                copied * sizeof(c8mi_lepton_transport_record));
    host_count_ += copied;
    statistics_.host_tracks_enqueued += copied;
    records += copied;
    count -= copied;
    if (host_count_ == kHostTrackBatchCapacity) flush_host_lepton_tracks();
  }
}

void RadioAccumulator::flush_host_lepton_tracks() {
  if (host_count_ == 0) return;
  C8MI_HIP(hipMemcpyAsync(host_batch_device_.data(), host_batch_.data(),
                          host_count_ * sizeof(c8mi_lepton_transport_record),
                          hipMemcpyHostToDevice, stream_));
  // The pinned batch is refilled by the next enqueue; finish its upload first
  // (RadioAccumulator.inl:277-281). Projection stays asynchronous.
  C8MI_HIP(hipStreamSynchronize(stream_));
  statistics_.host_to_device_bytes += host_count_ * sizeof(c8mi_lepton_transport_record);
  ++statistics_.host_track_batches;
  std::size_t const count = host_count_;
  host_count_ = 0;
  accumulate_lepton_records(host_batch_device_.data(), count);
}

void RadioAccumulator::ensure_track_capacity(std::uint64_t count) {
  if (track_capacity_ >= count) return;
  std::uint64_t const grown = track_capacity_ == 0 ? count : std::max(count, 2 * track_capacity_);
  if (track_capacity_ != 0) C8MI_HIP(hipStreamSynchronize(stream_));
  tracks_.allocate(C8MI_RADIO_TRACK_FIELDS * grown);
  track_capacity_ = grown;
}

void RadioAccumulator::accumulate_lepton_records(
    c8mi_lepton_transport_record const* device_records, std::size_t count) {
  if (count == 0) return;
  if (sealed_) throw std::logic_error("radio endpoint already committed; reset before new tracks");
  if (device_records == nullptr) throw std::invalid_argument("null device radio records");
  ensure_track_capacity(count);
  launch_records_to_tracks(device_records, count, tracks_.data(), track_capacity_, stream_);
  project(tracks_.data(), track_capacity_, count);
}

void RadioAccumulator::accumulate_tracks(double const* device_tracks, std::uint64_t stride,
                                         std::uint64_t count) {
  if (count == 0) return;
  if (sealed_) throw std::logic_error("radio endpoint already committed; reset before new tracks");
  if (device_tracks == nullptr) throw std::invalid_argument("null device radio tracks");
  if (stride < count) throw std::invalid_argument("radio track stride below track count");
  project(device_tracks, stride, count);
}

void RadioAccumulator::check_accumulate(std::uint64_t count) const {
  if (count == 0) return;
  if (sealed_) throw std::logic_error("radio endpoint already committed; reset before new tracks");
  // the launch shapes project() will build (they throw on oversized batches)
  bool const coreas = !coreas_host_.empty(), zhs = !zhs_host_.empty();
  if (fuse_ && !split_) {
    (void)base_launch(true, true, static_cast<std::uint32_t>(coreas_host_.size()),
                      std::max(coreas_max_bins_, zhs_max_bins_), count);
    return;
  }
  if (coreas)
    (void)base_launch(true, false, static_cast<std::uint32_t>(coreas_host_.size()), coreas_max_bins_, count);
  if (zhs) (void)base_launch(false, true, static_cast<std::uint32_t>(zhs_host_.size()), zhs_max_bins_, count);
}

ProjectionLaunch RadioAccumulator::base_launch(bool coreas, bool zhs, std::uint32_t observers,
                                               std::uint64_t max_bins,
                                               std::uint64_t count) const {
  ProjectionLaunch L{};
  L.observer_count = observers;
  L.zhs_refinement = config_.zhs_subtrack_refinement;
  L.propagation = propagation_;
  L.counters = counters_.data();
  LaunchTuning tuning;
  tuning.chunk_tracks = chunk_tracks_;
  tuning.tile_bins_limit = tile_bins_limit_;
  tuning.target_workgroups = target_workgroups_;
  tuning.block_threads = block_threads_;
  shape_projection(L, coreas, zhs, observers, max_bins, count, tuning);
  return L;
}

void RadioAccumulator::project(double const* tracks, std::uint64_t stride, std::uint64_t count) {
  auto run = [&](bool coreas, bool zhs, bool count_valid) {
    std::uint32_t const observers = static_cast<std::uint32_t>(
        coreas ? coreas_host_.size() : zhs_host_.size());
    if (observers == 0) return;
    std::uint64_t const max_bins = std::max(coreas ? coreas_max_bins_ : 0, zhs ? zhs_max_bins_ : 0);
    ProjectionLaunch L = base_launch(coreas, zhs, observers, max_bins, count);
    L.tracks = tracks;
    L.stride = stride;
    L.track_count = count;
    L.coreas_observers = coreas_observers_.data();
    L.zhs_observers = zhs_observers_.data();
    bool const fixed = config_.deterministic != 0;
    L.coreas_waves = fixed ? static_cast<void*>(coreas_fixed_.data())
                           : static_cast<void*>(coreas_floating_.data());
    L.zhs_waves = fixed ? static_cast<void*>(zhs_fixed_.data())
                        : static_cast<void*>(zhs_floating_.data());
    L.coreas_total_bins = coreas_bins_;
    L.zhs_total_bins = zhs_bins_;
    L.count_valid = count_valid ? 1u : 0u;
    std::pair<hipEvent_t, hipEvent_t> events{nullptr, nullptr};
    if (timing_) {
      if (!free_events_.empty()) {
        events = free_events_.back();
        free_events_.pop_back();
      } else {
        C8MI_HIP(hipEventCreate(&events.first));
        C8MI_HIP(hipEventCreate(&events.second));
      }
      C8MI_HIP(hipEventRecord(events.first, stream_));
    }
    launch_projection(L, coreas, zhs, fixed, config_.exact_fp64 != 0, block_threads_, stream_);
    if (timing_) {
      C8MI_HIP(hipEventRecord(events.second, stream_));
      pending_events_.push_back({events.first, events.second, coreas && zhs ? 0 : coreas ? 1 : 2});
    }
    ++statistics_.projection_launches;
    statistics_.track_observer_pairs += count * observers * (std::uint64_t(coreas) + zhs);
    if (coreas && zhs) statistics_.fused_track_observer_pairs += count * observers;
  };
  if (fuse_ && !split_) {
    run(true, true, true);
  } else {
    bool first = true;
    if (config_.coreas_enabled) {
      run(true, false, first);
      first = false;
    }
    if (config_.zhs_enabled) run(false, true, first);
  }
  downloaded_ = false;
}

void RadioAccumulator::copy_counters(c8mi_radio_counters* out) {
  CountersDev host{};
  C8MI_HIP(hipMemcpyAsync(&host, counters_.data(), sizeof(host), hipMemcpyDeviceToHost, stream_));
  C8MI_HIP(hipStreamSynchronize(stream_));
  statistics_.device_to_host_bytes += sizeof(host);
  if (out) {
    out->coreas_contributions = host.coreas_contributions;
    out->zhs_contributions = host.zhs_contributions;
    out->zhs_subtracks = host.zhs_subtracks;
    out->valid_tracks = host.valid_tracks;
    out->fixed_point_overflows = host.fixed_point_overflows;
  }
}

void RadioAccumulator::download(double* coreas, double* zhs, c8mi_radio_counters* counters) {
  if (downloaded_ || sealed_) throw std::logic_error("radio waveforms downloaded twice");
  flush_host_lepton_tracks();
  c8mi_radio_counters c{};
  copy_counters(&c);
  if (counters) *counters = c;
  if (c.fixed_point_overflows != 0)
    throw OverflowError(
        "radio fixed-point accumulator overflowed; increase fixed_point_field_limit_v_per_m");
  double const limit = config_.fixed_point_field_limit_v_per_m;
  auto fetch = [&](double* out, std::vector<c8mi_radio_observer> const& snapshots,
                   std::uint64_t total, DeviceBuffer<long long> const& fixed,
                   DeviceBuffer<double> const& floating, bool potential) {
    if (out == nullptr || total == 0) return;
    if (fixed.size()) {
      std::vector<long long> raw(3 * total);
      C8MI_HIP(hipMemcpyAsync(raw.data(), fixed.data(), fixed.bytes(), hipMemcpyDeviceToHost,
                              stream_));
      C8MI_HIP(hipStreamSynchronize(stream_));
      statistics_.device_to_host_bytes += fixed.bytes();
      // downloadSet (RadioAccumulator.inl:850-899).
      std::uint64_t offset = 0;
      for (auto const& o : snapshots) {
        double const factor = potential ? o.sample_rate_hz : 1.;
        double const inverse = limit / (kHeadroom * factor);
        for (std::uint64_t b = 0; b < o.number_of_bins; ++b) {
          std::uint64_t const j = offset + b;
          out[j] = static_cast<double>(raw[j]) * inverse;
          out[total + j] = static_cast<double>(raw[total + j]) * inverse;
          out[2 * total + j] = static_cast<double>(raw[2 * total + j]) * inverse;
        }
        offset += o.number_of_bins;
      }
    } else {
      C8MI_HIP(hipMemcpyAsync(out, floating.data(), floating.bytes(), hipMemcpyDeviceToHost,
                              stream_));
      C8MI_HIP(hipStreamSynchronize(stream_));
      statistics_.device_to_host_bytes += floating.bytes();
    }
  };
  fetch(coreas, coreas_snapshots_, coreas_bins_, coreas_fixed_, coreas_floating_, false);
  fetch(zhs, zhs_snapshots_, zhs_bins_, zhs_fixed_, zhs_floating_, true);
  downloaded_ = true;
}

void RadioAccumulator::download_fixed(std::int64_t* coreas, std::int64_t* zhs,
                                      c8mi_radio_counters* counters) {
  if (!config_.deterministic) throw std::logic_error("fixed radio snapshot needs deterministic mode");
  if (downloaded_ || sealed_) throw std::logic_error("fixed radio snapshot already committed");
  flush_host_lepton_tracks();
  sealed_ = true;
  copy_counters(counters);
  if (coreas && coreas_fixed_.size())
    C8MI_HIP(hipMemcpyAsync(coreas, coreas_fixed_.data(), coreas_fixed_.bytes(),
                            hipMemcpyDeviceToHost, stream_));
  if (zhs && zhs_fixed_.size())
    C8MI_HIP(hipMemcpyAsync(zhs, zhs_fixed_.data(), zhs_fixed_.bytes(), hipMemcpyDeviceToHost,
                            stream_));
  C8MI_HIP(hipStreamSynchronize(stream_));
  statistics_.device_to_host_bytes += coreas_fixed_.bytes() + zhs_fixed_.bytes();
  downloaded_ = true;
}

void RadioAccumulator::set_tuning(std::uint64_t chunk_tracks, std::uint32_t tile_bins,
                                  std::uint32_t target_workgroups, std::uint32_t block_threads,
                                  bool split) {
  std::uint32_t const block = block_threads == 0 ? kProjectionBlock : block_threads;
  if (block != 256 && block != 512 && block != 1024)
    throw std::invalid_argument("radio workgroup size must be 256, 512 or 1024");
  if (chunk_tracks % block != 0)
    throw std::invalid_argument("radio chunk size must be a multiple of the workgroup size");
  chunk_tracks_ = chunk_tracks;
  tile_bins_limit_ = tile_bins;
  if (target_workgroups != 0) target_workgroups_ = target_workgroups;
  block_threads_ = block;
  split_ = split;
}

void RadioAccumulator::enable_kernel_timing(bool enable) {
  timing_ = enable;
  for (int k = 0; k < 3; ++k) {
    timed_ms_[k] = 0.;
    timed_launches_[k] = 0;
  }
}

void RadioAccumulator::kernel_time_ms(double ms[3], std::uint64_t launches[3]) {
  C8MI_HIP(hipStreamSynchronize(stream_));
  for (auto& e : pending_events_) {
    float elapsed = 0.f;
    C8MI_HIP(hipEventElapsedTime(&elapsed, e.start, e.stop));
    timed_ms_[e.kind] += elapsed;
    ++timed_launches_[e.kind];
    free_events_.push_back({e.start, e.stop});
  }
  pending_events_.clear();
  for (int k = 0; k < 3; ++k) {
    if (ms) ms[k] = timed_ms_[k];
    if (launches) launches[k] = timed_launches_[k];
  }
}

}  // namespace c8mi::radio
