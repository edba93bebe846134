// Validation and device forms of radio configurations, shared by the radio
// accumulator and the per-slot radio of the EM pool.
#include "radio_setup.hpp"

#include "radio_launch.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace c8mi::radio {

// The observer checks of RadioAccumulator::validate: finite fields, positive
// rate, bins in [1, 2^31), positions inside the propagation table.
void validate_observers(c8mi_radio_observer const* observers, std::uint64_t count,
                        c8mi_radio_propagation const& p) {
  bool const homogeneous = p.homogeneous_refractive_index > 0.;
  if (count != 0 && observers == nullptr)
    throw std::invalid_argument("null radio observer list");
  for (std::uint64_t i = 0; i < count; ++i) {
    auto const& o = observers[i];
    if (!std::isfinite(o.start_time_s) || !std::isfinite(o.duration_s) ||
        !(o.duration_s >= 0.) || !std::isfinite(o.sample_rate_hz) ||
        !(o.sample_rate_hz > 0.) || o.number_of_bins == 0)
      throw std::invalid_argument("invalid radio observer");
    if (o.number_of_bins >= (std::uint64_t{1} << 31))
      throw std::length_error("radio observer with 2^31 or more bins");
    double const height = (o.position_m[2] - p.minimum_height_m) * p.inverse_step_per_m;
    if (!homogeneous && (!std::isfinite(height) || height < 0. ||
                         height + 0.5 >= static_cast<double>(p.table_size)))
      throw std::invalid_argument("radio observer outside propagation table");
    for (double coordinate : o.position_m)
      if (!std::isfinite(coordinate))
        throw std::invalid_argument("radio observer position is not finite");
  }
}

// RadioAccumulator::validate (corsika/detail/kokkos/radio/RadioAccumulator.inl:934-975).
void validate_radio_config(c8mi_radio_config const& config) {
  auto const& p = config.propagation;
  if (!config.coreas_enabled && !config.zhs_enabled)
    throw std::invalid_argument("enabled radio has no algorithm");
  if (!std::isfinite(config.fixed_point_field_limit_v_per_m) ||
      !(config.fixed_point_field_limit_v_per_m > 0.))
    throw std::invalid_argument("invalid radio fixed-point limit");
  if (config.zhs_subtrack_refinement < 1 || config.zhs_subtrack_refinement > 64)
    throw std::invalid_argument("ZHS subtrack refinement must be in [1,64]");
  if (!std::isfinite(p.homogeneous_refractive_index) || p.homogeneous_refractive_index < 0.)
    throw std::invalid_argument("invalid homogeneous radio refractive index");
  bool const homogeneous = p.homogeneous_refractive_index > 0.;
  if (!homogeneous &&
      (p.table_size < 11 || p.refractivity == nullptr || p.integrated_refractivity == nullptr ||
       !std::isfinite(p.minimum_height_m) || !std::isfinite(p.maximum_height_m) ||
       !(p.maximum_height_m > p.minimum_height_m) || !std::isfinite(p.step_m) ||
       !(p.step_m > 0.) || !std::isfinite(p.inverse_step_per_m) ||
       !(p.inverse_step_per_m > 0.)))
    throw std::invalid_argument("invalid flat-atmosphere radio snapshot");
  if (!homogeneous && p.table_size >= (std::uint64_t{1} << 32))
    throw std::length_error("radio propagation table above 2^32 entries");
  validate_observers(config.coreas_observers, config.coreas_enabled ? config.coreas_observer_count : 0, p);
  validate_observers(config.zhs_observers, config.zhs_enabled ? config.zhs_observer_count : 0, p);
}

// RadioAccumulator::makeObservers (RadioAccumulator.inl:591-617).
std::vector<ObserverDev> make_observers(c8mi_radio_observer const* input, std::uint64_t count,
                                        double limit, bool zhs_potential,
                                        std::uint64_t& total_bins, std::uint64_t& max_bins) {
  std::vector<ObserverDev> result;
  result.reserve(count);
  total_bins = 0;
  max_bins = 0;
  for (std::uint64_t i = 0; i < count; ++i) {
    auto const& o = input[i];
    ObserverDev d{};
    for (int a = 0; a < 3; ++a) d.position[a] = o.position_m[a];
    d.start_time = o.start_time_s;
    d.duration = o.duration_s;
    d.sample_rate = o.sample_rate_hz;
    double const potential_factor = zhs_potential ? o.sample_rate_hz : 1.;
    d.fixed_scale = kHeadroom * potential_factor / limit;
    d.number_of_bins = o.number_of_bins;
    d.waveform_offset = total_bins;
    if (o.number_of_bins > std::numeric_limits<std::uint64_t>::max() - total_bins)
      throw std::overflow_error("radio waveform bin overflow");
    total_bins += o.number_of_bins;
    max_bins = std::max<std::uint64_t>(max_bins, o.number_of_bins);
    result.push_back(d);
  }
  return result;
}

// observersCanShareTrackKinematics (RadioAccumulator.inl:619-632).
bool can_fuse(std::vector<ObserverDev> const& a, std::vector<ObserverDev> const& b) {
  if (a.size() != b.size() || a.empty()) return false;
  for (std::size_t i = 0; i < a.size(); ++i) {
    if (a[i].start_time != b[i].start_time || a[i].duration != b[i].duration ||
        a[i].sample_rate != b[i].sample_rate || a[i].number_of_bins != b[i].number_of_bins ||
        a[i].position[0] != b[i].position[0] || a[i].position[1] != b[i].position[1] ||
        a[i].position[2] != b[i].position[2])
      return false;
  }
  return true;
}


PropagationDev propagation_view(c8mi_radio_propagation const& p, double const* refractivity,
                                double const* integrated) {
  PropagationDev d{};
  d.homogeneous_index = p.homogeneous_refractive_index;
  d.minimum_height = p.minimum_height_m;
  d.maximum_height = p.maximum_height_m;
  d.step = p.step_m;
  d.inverse_step = p.inverse_step_per_m;
  d.slope_n_lower = p.slope_refractivity_lower;
  d.slope_i_lower = p.slope_integrated_refractivity_lower;
  d.slope_n_upper = p.slope_refractivity_upper;
  d.slope_i_upper = p.slope_integrated_refractivity_upper;
  d.refractivity = refractivity;
  d.integrated = integrated;
  d.table_size = p.homogeneous_refractive_index > 0. ? 0 : p.table_size;
  return d;
}

void shape_projection(ProjectionLaunch& L, bool coreas, bool zhs, std::uint32_t observers,
                      std::uint64_t max_bins, std::uint64_t count, LaunchTuning const& tuning) {
  // Window tile: the whole waveform when it fits the preferred LDS size, else
  // equal tiles of at most that size (a tuning limit may ask for others up to
  // the LDS budget).
  std::size_t const words = kMaxDynamicLdsBytes / sizeof(long long);
  std::size_t const per_bin = (coreas ? 3u : 0u) + (zhs ? 3u : 0u);
  std::uint64_t capacity = default_tile_capacity(per_bin, zhs);
  if (tuning.tile_bins_limit != 0)
    capacity = std::min<std::uint64_t>((words - 3) / per_bin, tuning.tile_bins_limit);
  std::uint64_t const want = std::max<std::uint64_t>(max_bins, 1);
  std::uint64_t tile = std::min<std::uint64_t>(want, capacity);
  if (tuning.tile_bins_limit == 0) {  // equal tiles: a 4096-bin window as two 2048-bin tiles
    std::uint64_t const tiles = (want + capacity - 1) / capacity;
    tile = (want + tiles - 1) / tiles;
  }
  L.tile_bins = static_cast<std::uint32_t>(tile);
  L.bin_tiles = static_cast<std::uint32_t>((max_bins + tile - 1) / tile);
  if (L.bin_tiles == 0) L.bin_tiles = 1;
  std::uint64_t chunk = tuning.chunk_tracks;
  if (chunk == 0) {
    std::uint64_t const per_chunk = std::max<std::uint64_t>(1, std::uint64_t(observers) * L.bin_tiles);
    std::uint64_t const chunks_wanted = std::max<std::uint64_t>(1, tuning.target_workgroups / per_chunk);
    // a workgroup clears and flushes a whole window tile, so chunks of tens
    // of thousands of tracks pay (2^20 tracks on 160 antennas: 32768-track
    // chunks 1.2x faster than 16384, as fast as 65536)
    std::uint64_t chunks = std::min<std::uint64_t>(chunks_wanted, count / kMinimumDefaultChunk);
    // chunk c runs on XCD c % 8: a multiple of eight chunks keeps the XCDs
    // evenly loaded (50 chunks measured 12% slower than 48 or 64); a small
    // batch (a lepton front's tracks) still gets eight chunks, so every XCD
    // works on it (the EM + radio shower measured 10% faster than with
    // chunks of at least 8192 tracks)
    chunks = chunks >= kXcdCount ? chunks / kXcdCount * kXcdCount : kXcdCount;
    chunk = (count + chunks - 1) / chunks;
    std::uint64_t const block = tuning.block_threads;
    chunk = std::max<std::uint64_t>(block, (chunk + block - 1) / block * block);
  }
  L.chunk_size = chunk;
  std::uint64_t const chunks = (count + chunk - 1) / chunk;
  if (chunks > std::numeric_limits<std::uint32_t>::max())
    throw std::length_error("radio track batch too large for one launch");
  L.chunks = static_cast<std::uint32_t>(chunks);
  std::uint64_t const grid = (chunks + 7) / 8 * 8 * std::uint64_t(observers) * L.bin_tiles;
  if (grid > std::numeric_limits<std::uint32_t>::max())
    throw std::length_error("radio projection grid too large");
}

}  // namespace c8mi::radio
