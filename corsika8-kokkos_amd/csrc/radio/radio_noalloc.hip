// One-shot radio projection that allocates no device memory.
//
// The device copies a projection needs beyond the caller's tracks and
// waveforms -- the observer lists, the propagation table and the counters --
// live in `scratch`, one file-scope device array of fixed size. Every GPU the
// library runs on holds its own copy of the array until the process ends; its
// address on each device is taken once, on the first call with that device
// current, through hipGetSymbolAddress, and only that address is handed to
// copies and kernels. Calls on one GPU share the array: the call clears what
// must start at zero, and calls that overlap on one GPU (other threads or
// streams) race.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <vector>

#include "c8mi/c8mi_radio.h"
#include "../runtime.hpp"
#include "radio_launch.hpp"
#include "radio_noalloc.hpp"
#include "radio_setup.hpp"

namespace c8mi::radio {
namespace {

constexpr std::size_t kCountersBytes = 256;
constexpr std::size_t kObserverBytes = C8MI_RADIO_NOALLOC_MAX_OBSERVERS * sizeof(ObserverDev);
constexpr std::size_t kTableBytes = C8MI_RADIO_NOALLOC_MAX_TABLE_ENTRIES * sizeof(double);
// counters | CoREAS observers | ZHS observers | refractivity | integrated
constexpr std::size_t kCoreasOffset = kCountersBytes;
constexpr std::size_t kZhsOffset = kCoreasOffset + kObserverBytes;
constexpr std::size_t kRefractivityOffset = kZhsOffset + kObserverBytes;
constexpr std::size_t kIntegratedOffset = kRefractivityOffset + kTableBytes;
constexpr std::size_t kScratchBytes = kIntegratedOffset + kTableBytes;
static_assert(sizeof(CountersDev) <= kCountersBytes, "counters fit their scratch slot");
static_assert(kObserverBytes % 256 == 0 && kTableBytes % 256 == 0, "256-byte aligned regions");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device (taken once per
// device; the array's host-side name is not a device address).
unsigned char* scratch_address() {
  int device = 0;
  C8MI_HIP(hipGetDevice(&device));
  if (device < 0 || device >= kMaxDevices)
    throw std::runtime_error("radio no-alloc projection: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch[device] == nullptr) {
    void* ptr = nullptr;
    C8MI_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (ptr == nullptr) throw DeviceError("radio no-alloc scratch has no device address");
    g_scratch[device] = static_cast<unsigned char*>(ptr);
  }
  return g_scratch[device];
}

__global__ void zero_words_kernel(unsigned long long* words, std::uint64_t count) {
  std::uint64_t const stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
  for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       i < count; i += stride)
    words[i] = 0ull;
}

// A fill kernel rather than hipMemsetAsync, so the call stays a plain sequence
// of kernels and copies on the caller's stream.
void zero_words(void* words, std::uint64_t count, hipStream_t stream) {
  if (count == 0) return;
  std::uint64_t const blocks = std::min<std::uint64_t>((count + 255) / 256, 4096);
  hipLaunchKernelGGL(zero_words_kernel, dim3(static_cast<std::uint32_t>(blocks)), dim3(256), 0,
                     stream, static_cast<unsigned long long*>(words), count);
  C8MI_HIP(hipGetLastError());
}

}  // namespace

void project_fixed_noalloc(c8mi_radio_config const& config, double const* device_tracks,
                           std::uint64_t stride, std::uint64_t count, std::int64_t* device_coreas,
                           std::int64_t* device_zhs, c8mi_radio_counters* counters,
                           hipStream_t stream) {
  validate_radio_config(config);
  if (!config.deterministic)
    throw std::invalid_argument("radio no-alloc projection writes int64 fixed point: deterministic must be 1");
  if (count != 0 && device_tracks == nullptr) throw std::invalid_argument("null device radio tracks");
  if (stride < count) throw std::invalid_argument("radio track stride below track count");
  bool const coreas = config.coreas_enabled && config.coreas_observer_count != 0;
  bool const zhs = config.zhs_enabled && config.zhs_observer_count != 0;
  if ((coreas && config.coreas_observer_count > C8MI_RADIO_NOALLOC_MAX_OBSERVERS) ||
      (zhs && config.zhs_observer_count > C8MI_RADIO_NOALLOC_MAX_OBSERVERS))
    throw std::length_error("radio no-alloc projection: more observers than its scratch holds (1024 per algorithm)");
  auto const& p = config.propagation;
  bool const tabulated = !(p.homogeneous_refractive_index > 0.);
  if (tabulated && p.table_size > C8MI_RADIO_NOALLOC_MAX_TABLE_ENTRIES)
    throw std::length_error("radio no-alloc projection: propagation table longer than its scratch holds (131072 entries)");
  if ((coreas && device_coreas == nullptr) || (zhs && device_zhs == nullptr))
    throw std::invalid_argument("null device waveform output for an enabled algorithm");

  double const limit = config.fixed_point_field_limit_v_per_m;
  std::uint64_t coreas_bins = 0, coreas_max = 0, zhs_bins = 0, zhs_max = 0;
  std::vector<ObserverDev> coreas_host, zhs_host;
  if (coreas)
    coreas_host = make_observers(config.coreas_observers, config.coreas_observer_count, limit,
                                 false, coreas_bins, coreas_max);
  if (zhs)
    zhs_host = make_observers(config.zhs_observers, config.zhs_observer_count, limit, true,
                              zhs_bins, zhs_max);

  unsigned char* const base = scratch_address();
  auto* const device_counters = reinterpret_cast<CountersDev*>(base);
  auto* const coreas_observers = reinterpret_cast<ObserverDev*>(base + kCoreasOffset);
  auto* const zhs_observers = reinterpret_cast<ObserverDev*>(base + kZhsOffset);
  auto* const refractivity = reinterpret_cast<double*>(base + kRefractivityOffset);
  auto* const integrated = reinterpret_cast<double*>(base + kIntegratedOffset);

  // Outputs and counters start at zero; the previous call's scratch contents
  // are overwritten before any kernel reads them.
  zero_words(device_counters, sizeof(CountersDev) / sizeof(unsigned long long), stream);
  if (coreas) zero_words(device_coreas, 3 * coreas_bins, stream);
  if (zhs) zero_words(device_zhs, 3 * zhs_bins, stream);
  if (coreas)
    C8MI_HIP(hipMemcpyAsync(coreas_observers, coreas_host.data(), coreas_host.size() * sizeof(ObserverDev),
                            hipMemcpyHostToDevice, stream));
  if (zhs)
    C8MI_HIP(hipMemcpyAsync(zhs_observers, zhs_host.data(), zhs_host.size() * sizeof(ObserverDev),
                            hipMemcpyHostToDevice, stream));
  if (tabulated) {
    C8MI_HIP(hipMemcpyAsync(refractivity, p.refractivity, p.table_size * sizeof(double),
                            hipMemcpyHostToDevice, stream));
    C8MI_HIP(hipMemcpyAsync(integrated, p.integrated_refractivity, p.table_size * sizeof(double),
                            hipMemcpyHostToDevice, stream));
  }

  // The accumulator's default launch shape: CoREAS and ZHS as separate
  // launches, valid tracks tallied by the first.
  LaunchTuning const tuning;
  bool first = true;
  auto run = [&](bool c, bool z) {
    std::uint32_t const observers = static_cast<std::uint32_t>(c ? coreas_host.size() : zhs_host.size());
    ProjectionLaunch L{};
    L.observer_count = observers;
    L.zhs_refinement = config.zhs_subtrack_refinement;
    L.propagation = propagation_view(p, refractivity, integrated);
    L.counters = device_counters;
    shape_projection(L, c, z, observers, c ? coreas_max : zhs_max, count, tuning);
    L.tracks = device_tracks;
    L.stride = stride;
    L.track_count = count;
    L.coreas_observers = coreas_observers;
    L.zhs_observers = zhs_observers;
    L.coreas_waves = device_coreas;
    L.zhs_waves = device_zhs;
    L.coreas_total_bins = coreas_bins;
    L.zhs_total_bins = zhs_bins;
    L.count_valid = first ? 1u : 0u;
    first = false;
    launch_projection(L, c, z, true, config.exact_fp64 != 0, tuning.block_threads, stream);
  };
  if (count != 0) {
    if (coreas) run(true, false);
    if (zhs) run(false, true);
  }

  CountersDev host{};
  C8MI_HIP(hipMemcpyAsync(&host, device_counters, sizeof(host), hipMemcpyDeviceToHost, stream));
  C8MI_HIP(hipStreamSynchronize(stream));
  if (counters) {
    counters->coreas_contributions = host.coreas_contributions;
    counters->zhs_contributions = host.zhs_contributions;
    counters->zhs_subtracks = host.zhs_subtracks;
    counters->valid_tracks = host.valid_tracks;
    counters->fixed_point_overflows = host.fixed_point_overflows;
  }
  if (host.fixed_point_overflows != 0)
    throw OverflowError(
        "radio fixed-point accumulator overflowed; increase fixed_point_field_limit_v_per_m");
}

}  // namespace c8mi::radio
