// Host side of the per-shower radio of a shower pool. See slot_radio.hpp.
#include "slot_radio.hpp"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

#include "radio_launch.hpp"
#include "radio_setup.hpp"

namespace c8mi::radio {

SlotRadio::SlotRadio(c8mi_radio_config const& config, std::uint32_t slots, std::uint32_t observers_per_slot,
                     std::uint32_t bins_per_observer, hipStream_t stream)
    : config_(config), slots_(slots), observers_per_slot_(observers_per_slot), bins_per_observer_(bins_per_observer) {
  c8mi_radio_config check = config;
  check.coreas_observers = check.zhs_observers = nullptr;
  check.coreas_observer_count = check.zhs_observer_count = 0;
  validate_radio_config(check);
  if (!config.deterministic) throw std::invalid_argument("per-shower radio accumulates in int64 fixed point only");
  if (slots == 0 || observers_per_slot == 0 || observers_per_slot > 256)
    throw std::invalid_argument("per-shower radio needs 1..256 observers per slot");
  if (bins_per_observer == 0 || bins_per_observer > (1u << 20))
    throw std::invalid_argument("per-shower radio bins per observer must be in [1, 2^20]");
  config_.coreas_observers = config_.zhs_observers = nullptr;
  config_.coreas_observer_count = config_.zhs_observer_count = 0;
  total_bins_ = static_cast<std::uint64_t>(slots) * observers_per_slot * bins_per_observer;
  // Both algorithms share every observer, so they share one launch.
  fuse_ = config.coreas_enabled && config.zhs_enabled;
  std::size_t const per_bin = (config.coreas_enabled ? 3u : 0u) + (config.zhs_enabled ? 3u : 0u);
  std::uint64_t const capacity = default_tile_capacity(per_bin, config.zhs_enabled != 0);
  std::uint64_t const tiles = (bins_per_observer + capacity - 1) / capacity;
  tile_bins_ = static_cast<std::uint32_t>((bins_per_observer + tiles - 1) / tiles);
  bin_tiles_ = static_cast<std::uint32_t>((bins_per_observer + tile_bins_ - 1) / tile_bins_);
  if (stream == nullptr) {
    C8MI_HIP(hipStreamCreate(&stream_));
    owns_stream_ = true;
  } else {
    stream_ = stream;
  }
  try {
    auto const& p = config.propagation;
    if (!(p.homogeneous_refractive_index > 0.)) {
      refractivity_.allocate(p.table_size);
      integrated_.allocate(p.table_size);
      C8MI_HIP(hipMemcpy(refractivity_.data(), p.refractivity, p.table_size * sizeof(double),
                         hipMemcpyHostToDevice));
      C8MI_HIP(hipMemcpy(integrated_.data(), p.integrated_refractivity, p.table_size * sizeof(double),
                         hipMemcpyHostToDevice));
    }
    propagation_ = propagation_view(p, refractivity_.data(), integrated_.data());
    config_.propagation.refractivity = config_.propagation.integrated_refractivity = nullptr;
    std::size_t const observers = static_cast<std::size_t>(slots) * observers_per_slot;
    coreas_observers_.allocate(observers);
    zhs_observers_.allocate(observers);
    observer_counts_.allocate(slots);
    C8MI_HIP(hipMemset(observer_counts_.data(), 0, observer_counts_.bytes()));
    if (config.coreas_enabled) coreas_fixed_.allocate(3 * total_bins_);
    if (config.zhs_enabled) zhs_fixed_.allocate(3 * total_bins_);
    counters_.allocate(slots);
    C8MI_HIP(hipMemset(counters_.data(), 0, counters_.bytes()));
    slot_observers_.resize(slots);
    counts_host_.assign(slots, 0);
    staging_.resize(observers_per_slot);
  } catch (...) {
    if (owns_stream_) (void)hipStreamDestroy(stream_);
    throw;
  }
}

SlotRadio::~SlotRadio() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& e : pending_events_)
    for (auto ev : e) (void)hipEventDestroy(ev);
  if (owns_stream_) (void)hipStreamDestroy(stream_);
}

std::uint64_t SlotRadio::slot_bins(std::uint32_t slot) const {
  std::uint64_t bins = 0;
  for (auto const& o : slot_observers_.at(slot)) bins += o.number_of_bins;
  return bins;
}

void SlotRadio::set_slot(std::uint32_t slot, c8mi_radio_observer const* observers, std::uint64_t count) {
  if (slot >= slots_) throw std::invalid_argument("radio slot outside the pool");
  if (count > observers_per_slot_) throw std::invalid_argument("more radio observers than a slot holds");
  if (count && !observers) throw std::invalid_argument("null radio observer list");
  for (std::uint64_t k = 0; k < count; ++k)
    if (observers[k].number_of_bins > bins_per_observer_)
      throw std::invalid_argument("radio observer with more bins than a slot's observer holds");
  validate_observers(observers, count, config_.propagation);

  double const limit = config_.fixed_point_field_limit_v_per_m;
  std::uint64_t const first = static_cast<std::uint64_t>(slot) * observers_per_slot_;
  auto fill = [&](bool potential, DeviceBuffer<ObserverDev>& device) {
    std::uint64_t total = 0, max_bins = 0;
    auto made = make_observers(observers, count, limit, potential, total, max_bins);
    for (std::uint64_t k = 0; k < count; ++k) {
      made[k].waveform_offset = (first + k) * bins_per_observer_;
      staging_[k] = made[k];
    }
    if (count)
      C8MI_HIP(hipMemcpyAsync(device.data() + first, staging_.data(), count * sizeof(ObserverDev),
                              hipMemcpyHostToDevice, stream_));
    C8MI_HIP(hipStreamSynchronize(stream_));  // the host staging is reused
  };
  fill(false, coreas_observers_);
  fill(true, zhs_observers_);
  std::uint32_t const n = static_cast<std::uint32_t>(count);
  C8MI_HIP(hipMemcpyAsync(observer_counts_.data() + slot, &n, sizeof n, hipMemcpyHostToDevice, stream_));
  std::uint64_t const region = static_cast<std::uint64_t>(observers_per_slot_) * bins_per_observer_;
  for (DeviceBuffer<long long>* w : {&coreas_fixed_, &zhs_fixed_}) {
    if (!w->size()) continue;
    for (int c = 0; c < 3; ++c)
      C8MI_HIP(hipMemsetAsync(w->data() + c * total_bins_ + first * bins_per_observer_, 0,
                              region * sizeof(long long), stream_));
  }
  C8MI_HIP(hipMemsetAsync(counters_.data() + slot, 0, sizeof(CountersDev), stream_));
  C8MI_HIP(hipStreamSynchronize(stream_));
  slot_observers_[slot].assign(observers, observers + count);
  counts_host_[slot] = n;
}

void SlotRadio::download(std::uint32_t slot, std::int64_t* coreas, std::int64_t* zhs,
                         c8mi_radio_counters* counters) {
  if (slot >= slots_) throw std::invalid_argument("radio slot outside the pool");
  std::uint64_t const first = static_cast<std::uint64_t>(slot) * observers_per_slot_;
  std::uint64_t const region = static_cast<std::uint64_t>(observers_per_slot_) * bins_per_observer_;
  auto const& obs = slot_observers_[slot];
  std::uint64_t const bins = slot_bins(slot);
  download_staging_.resize(3 * region);
  auto fetch = [&](DeviceBuffer<long long> const& w, std::int64_t* out) {
    if (!out || !w.size()) return;
    for (int c = 0; c < 3; ++c)
      C8MI_HIP(hipMemcpyAsync(download_staging_.data() + c * region,
                              w.data() + c * total_bins_ + first * bins_per_observer_, region * sizeof(long long),
                              hipMemcpyDeviceToHost, stream_));
    C8MI_HIP(hipStreamSynchronize(stream_));
    for (int c = 0; c < 3; ++c) {
      std::uint64_t at = 0;
      for (std::size_t k = 0; k < obs.size(); ++k) {
        std::memcpy(out + c * bins + at, download_staging_.data() + c * region + k * bins_per_observer_,
                    obs[k].number_of_bins * sizeof(long long));
        at += obs[k].number_of_bins;
      }
    }
  };
  fetch(coreas_fixed_, coreas);
  fetch(zhs_fixed_, zhs);
  CountersDev host{};
  C8MI_HIP(hipMemcpyAsync(&host, counters_.data() + slot, sizeof host, hipMemcpyDeviceToHost, stream_));
  C8MI_HIP(hipStreamSynchronize(stream_));
  if (counters) {
    counters->coreas_contributions = host.coreas_contributions;
    counters->zhs_contributions = host.zhs_contributions;
    counters->zhs_subtracks = host.zhs_subtracks;
    counters->valid_tracks = host.valid_tracks;
    counters->fixed_point_overflows = host.fixed_point_overflows;
  }
}

// Every chunk of every segment on every observer: at most one partial chunk
// per segment beyond count / chunk.
std::uint64_t SlotRadio::grid_for(std::uint64_t count, std::uint64_t chunk, std::uint64_t max_segments) const {
  return ((count + chunk - 1) / chunk + max_segments) * static_cast<std::uint64_t>(observers_per_slot_) * bin_tiles_;
}

void SlotRadio::check_project(std::uint64_t count, std::uint64_t chunk, std::uint64_t max_segments) const {
  if (count == 0) return;
  if (chunk == 0) throw std::invalid_argument("per-shower radio chunk of zero tracks");
  if (grid_for(count, chunk, max_segments) > 0x7fffffffULL)
    throw std::length_error("per-shower radio launch above the grid limit");
}

std::uint64_t SlotRadio::chunk_for(std::uint64_t count) const noexcept {
  constexpr std::uint64_t kTargetItems = 8192, kMinimumChunk = 4096, kMaximumChunk = 1 << 20;
  std::uint64_t const per_item = static_cast<std::uint64_t>(observers_per_slot_) * bin_tiles_;
  std::uint64_t chunk = (count * per_item + kTargetItems - 1) / kTargetItems;
  return std::min(kMaximumChunk, std::max(kMinimumChunk, chunk));
}

void SlotRadio::project(double const* tracks, std::uint64_t stride, std::uint64_t count, std::uint64_t chunk,
                        SegmentDev const* segments, std::uint32_t const* segment_count,
                        unsigned long long const* item_count, std::uint32_t max_segments) {
  if (count == 0) return;
  ProjectionLaunch L{};
  L.tracks = tracks;
  L.stride = stride;
  L.track_count = count;
  L.coreas_observers = coreas_observers_.data();
  L.zhs_observers = zhs_observers_.data();
  L.observer_count = observers_per_slot_;
  L.zhs_refinement = config_.zhs_subtrack_refinement;
  L.propagation = propagation_;
  L.coreas_waves = coreas_fixed_.data();
  L.zhs_waves = zhs_fixed_.data();
  L.coreas_total_bins = total_bins_;
  L.zhs_total_bins = total_bins_;
  L.counters = counters_.data();
  L.chunk_size = chunk;
  L.tile_bins = tile_bins_;
  L.bin_tiles = bin_tiles_;
  L.segments = segments;
  L.segment_count = segment_count;
  L.item_count = item_count;
  L.slot_observer_counts = observer_counts_.data();
  L.slot_observers = observers_per_slot_;
  check_project(count, chunk, max_segments);
  std::uint64_t const grid = grid_for(count, chunk, max_segments);
  std::array<hipEvent_t, 2> events{nullptr, nullptr};
  if (timing_) {
    C8MI_HIP(hipEventCreate(&events[0]));
    C8MI_HIP(hipEventCreate(&events[1]));
    C8MI_HIP(hipEventRecord(events[0], stream_));
  }
  bool const exact = config_.exact_fp64 != 0;
  auto const block = kProjectionBlock;
  auto const g = static_cast<std::uint32_t>(grid);
  if (fuse_) {
    launch_pooled_projection(L, true, true, exact, block, g, stream_);
  } else {
    if (config_.coreas_enabled) launch_pooled_projection(L, true, false, exact, block, g, stream_);
    if (config_.zhs_enabled) launch_pooled_projection(L, false, true, exact, block, g, stream_);
  }
  if (timing_) {
    C8MI_HIP(hipEventRecord(events[1], stream_));
    pending_events_.push_back(events);
  }
}

void SlotRadio::enable_kernel_timing(bool enable) {
  timing_ = enable;
  ms_ = 0.;
  launches_ = 0;
}

void SlotRadio::kernel_time_ms(double& ms, std::uint64_t& launches) {
  C8MI_HIP(hipStreamSynchronize(stream_));
  for (auto& e : pending_events_) {
    float elapsed = 0.f;
    C8MI_HIP(hipEventElapsedTime(&elapsed, e[0], e[1]));
    ms_ += elapsed;
    ++launches_;
    for (auto ev : e) C8MI_HIP(hipEventDestroy(ev));
  }
  pending_events_.clear();
  ms = ms_;
  launches = launches_;
}

}  // namespace c8mi::radio
