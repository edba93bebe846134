// extern "C" radio entry points (include/c8mi/c8mi_radio.h) over
// c8mi::radio::RadioAccumulator. Exceptions never cross the ABI: each is mapped
// to a status code and its message kept per thread.
#include <hip/hip_runtime.h>

#include <cmath>
#include <new>
#include <stdexcept>
#include <string>

#include "c8mi/c8mi_radio.h"
#include "radio/radio_accumulator.hpp"
#include "radio/radio_noalloc.hpp"
#include "capi_guard.hpp"
#include "capi_handles.hpp"
#include "runtime.hpp"

namespace c8mi {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(std::string message) { g_last_error = std::move(message); }
char const* last_error() { return g_last_error.c_str(); }

}  // namespace c8mi

using c8mi::guarded;

extern "C" {

char const* c8mi_last_error(void) { return c8mi::last_error(); }

int c8mi_radio_create(c8mi_radio_config const* config, void* stream, c8mi_radio** out) {
  return guarded([&] {
    if (config == nullptr || out == nullptr) throw std::invalid_argument("null argument");
    *out = nullptr;
    *out = new c8mi_radio(*config, static_cast<hipStream_t>(stream));
  });
}

int c8mi_radio_destroy(c8mi_radio* radio) {
  return guarded([&] { delete radio; });
}

int c8mi_radio_reset(c8mi_radio* radio) {
  return guarded([&] {
    if (!radio) throw std::invalid_argument("null radio handle");
    radio->accumulator.reset();
  });
}

int c8mi_radio_enqueue_host_tracks(c8mi_radio* radio, c8mi_lepton_transport_record const* records,
                                   uint64_t count) {
  return guarded([&] {
    if (!radio) throw std::invalid_argument("null radio handle");
    radio->accumulator.enqueue_host_lepton_tracks(records, count);
  });
}

int c8mi_radio_flush_host_tracks(c8mi_radio* radio) {
  return guarded([&] {
    if (!radio) throw std::invalid_argument("null radio handle");
    radio->accumulator.flush_host_lepton_tracks();
  });
}

int c8mi_radio_accumulate_records(c8mi_radio* radio,
                                  c8mi_lepton_transport_record const* device_records,
                                  uint64_t count) {
  return guarded([&] {
    if (!radio) throw std::invalid_argument("null radio handle");
    radio->accumulator.accumulate_lepton_records(device_records, count);
  });
}

int c8mi_radio_accumulate_tracks(c8mi_radio* radio, double const* device_tracks, uint64_t stride,
                                 uint64_t count) {
  return guarded([&] {
    if (!radio) throw std::invalid_argument("null radio handle");
    radio->accumulator.accumulate_tracks(device_tracks, stride, count);
  });
}

int c8mi_radio_bins(c8mi_radio const* radio, uint64_t* coreas_total_bins,
                    uint64_t* zhs_total_bins, int* fused) {
  return guarded([&] {
    if (!radio) throw std::invalid_argument("null radio handle");
    if (coreas_total_bins) *coreas_total_bins = radio->accumulator.coreas_total_bins();
    if (zhs_total_bins) *zhs_total_bins = radio->accumulator.zhs_total_bins();
    if (fused) *fused = radio->accumulator.fused() ? 1 : 0;
  });
}

int c8mi_radio_download(c8mi_radio* radio, double* coreas, double* zhs,
                        c8mi_radio_counters* counters) {
  return guarded([&] {
    if (!radio) throw std::invalid_argument("null radio handle");
    radio->accumulator.download(coreas, zhs, counters);
  });
}

int c8mi_radio_download_fixed(c8mi_radio* radio, int64_t* coreas, int64_t* zhs,
                              c8mi_radio_counters* counters) {
  return guarded([&] {
    if (!radio) throw std::invalid_argument("null radio handle");
    radio->accumulator.download_fixed(coreas, zhs, counters);
  });
}

int c8mi_radio_zhs_electric_field(c8mi_radio_observer const* observers, uint64_t observer_count,
                                  double const* potential, double* field, double* times) {
  return guarded([&] {
    if ((!observers && observer_count) || !potential || !field)
      throw std::invalid_argument("null argument");
    std::uint64_t total = 0;
    for (std::uint64_t k = 0; k < observer_count; ++k) {
      c8mi_radio_observer const& o = observers[k];
      if (o.number_of_bins < 2) throw std::invalid_argument("every observer needs at least two bins");
      if (!(std::isfinite(o.sample_rate_hz) && o.sample_rate_hz > 0.) || !std::isfinite(o.start_time_s))
        throw std::invalid_argument("observer sample rate and start time must be finite (rate > 0)");
      total += o.number_of_bins;
    }
    std::uint64_t const out_total = total - observer_count;
    std::uint64_t in_at = 0, out_at = 0;
    for (std::uint64_t k = 0; k < observer_count; ++k) {
      c8mi_radio_observer const& o = observers[k];
      std::uint64_t const bins = o.number_of_bins;
      double const rate = o.sample_rate_hz;
      long double const period = 1.0L / static_cast<long double>(rate);
      for (int c = 0; c < 3; ++c) {
        double const* a = potential + c * total + in_at;
        double* e = field + c * out_total + out_at;
        for (std::uint64_t i = 0; i + 1 < bins; ++i) e[i] = -(a[i + 1] - a[i]) * rate;
      }
      if (times) {
        long double const start = o.start_time_s;
        for (std::uint64_t i = 0; i + 1 < bins; ++i) {
          long double const t0 = start + static_cast<long double>(i) * period;
          long double const t1 = start + static_cast<long double>(i + 1) * period;
          times[out_at + i] = static_cast<double>((t0 + t1) / 2.0L);
        }
      }
      in_at += bins;
      out_at += bins - 1;
    }
  });
}

int c8mi_radio_set_tuning(c8mi_radio* radio, uint64_t chunk_tracks, uint32_t tile_bins,
                          uint32_t target_workgroups, uint32_t block_threads, int split) {
  return guarded([&] {
    if (!radio) throw std::invalid_argument("null radio handle");
    radio->accumulator.set_tuning(chunk_tracks, tile_bins, target_workgroups, block_threads,
                                  split != 0);
  });
}

int c8mi_radio_enable_kernel_timing(c8mi_radio* radio, int enable) {
  return guarded([&] {
    if (!radio) throw std::invalid_argument("null radio handle");
    radio->accumulator.enable_kernel_timing(enable != 0);
  });
}

int c8mi_radio_kernel_time(c8mi_radio* radio, double ms[3], uint64_t launches[3]) {
  return guarded([&] {
    if (!radio) throw std::invalid_argument("null radio handle");
    radio->accumulator.kernel_time_ms(ms, launches);
  });
}

int c8mi_radio_project_fixed_noalloc(c8mi_radio_config const* config, double const* device_tracks,
                                     uint64_t stride, uint64_t count, int64_t* device_coreas,
                                     int64_t* device_zhs, c8mi_radio_counters* counters,
                                     void* stream) {
  return guarded([&] {
    if (config == nullptr) throw std::invalid_argument("null radio config");
    c8mi::radio::project_fixed_noalloc(*config, device_tracks, stride, count, device_coreas,
                                       device_zhs, counters, static_cast<hipStream_t>(stream));
  });
}

int c8mi_radio_build_flat_table(double const* refractivity, uint64_t count,
                                double minimum_height_m, double maximum_height_m, double step_m,
                                double* integrated, c8mi_radio_propagation* out) {
  return guarded([&] {
    if (!refractivity || !integrated || !out) throw std::invalid_argument("null argument");
    if (count < 11) throw std::invalid_argument("radio propagation table needs eleven points");
    if (!(step_m > 0.) || !std::isfinite(step_m))
      throw std::invalid_argument("radio propagation table requires a positive step");
    double const inverse_step = 1. / step_m;
    // FlatAtmosphereConfig.inl:58-73: rectangular running sum and ten-bin slopes.
    integrated[0] = refractivity[0] * inverse_step;
    for (uint64_t i = 1; i < count; ++i)
      integrated[i] = integrated[i - 1] + refractivity[i] * inverse_step;
    uint64_t const last = count - 1;
    *out = c8mi_radio_propagation{};
    out->minimum_height_m = minimum_height_m;
    out->maximum_height_m = maximum_height_m;
    out->step_m = step_m;
    out->inverse_step_per_m = inverse_step;
    out->slope_refractivity_lower = (refractivity[10] - refractivity[0]) / 10.;
    out->slope_integrated_refractivity_lower = (integrated[10] - integrated[0]) / 10.;
    out->slope_refractivity_upper = (refractivity[last] - refractivity[last - 10]) / 10.;
    out->slope_integrated_refractivity_upper = (integrated[last] - integrated[last - 10]) / 10.;
    out->refractivity = refractivity;
    out->integrated_refractivity = integrated;
    out->table_size = count;
  });
}

int c8mi_device_count(int* count) {
  return guarded([&] {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    if (count) *count = n;
  });
}

int c8mi_device_malloc(void** pointer, uint64_t bytes) {
  return guarded([&] {
    if (!pointer) throw std::invalid_argument("null argument");
    *pointer = nullptr;
    if (bytes) C8MI_HIP(hipMalloc(pointer, bytes));
  });
}

int c8mi_device_free(void* pointer) {
  return guarded([&] {
    if (pointer) C8MI_HIP(hipFree(pointer));
  });
}

int c8mi_memcpy_htod(void* device_destination, void const* host_source, uint64_t bytes) {
  return guarded([&] {
    if (bytes) C8MI_HIP(hipMemcpy(device_destination, host_source, bytes, hipMemcpyHostToDevice));
  });
}

int c8mi_memcpy_dtoh(void* host_destination, void const* device_source, uint64_t bytes) {
  return guarded([&] {
    if (bytes) C8MI_HIP(hipMemcpy(host_destination, device_source, bytes, hipMemcpyDeviceToHost));
  });
}

int c8mi_memcpy_dtod(void* device_destination, void const* device_source, uint64_t bytes) {
  return guarded([&] {
    if (bytes)
      C8MI_HIP(hipMemcpy(device_destination, device_source, bytes, hipMemcpyDeviceToDevice));
  });
}

int c8mi_device_synchronize(void) {
  return guarded([&] { C8MI_HIP(hipDeviceSynchronize()); });
}

}  // extern "C"
