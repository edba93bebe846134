// Host side of the EM backend: device storage of tables, environment and
// context; per-species double-buffered particle queues; the front loop.
// Mirrors the reference's IEmBackend implementation
// (corsika/kokkos/em/execution/IEmBackend.hpp:25-88,
// corsika/detail/kokkos/em/execution/BackendInstanceImplementation.inl) with
// one HIP stream per backend and one small device-to-host read per front.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <memory>
#include <vector>

#include "c8mi/c8mi_em.h"
#include "em/device_tables.hpp"
#include "em/em_device.hpp"
#include "runtime.hpp"

namespace c8mi::em {

class QueueStorage {
 public:
  explicit QueueStorage(std::uint64_t capacity);
  ParticleQueue view() const noexcept { return view_; }
  std::uint64_t capacity() const noexcept { return capacity_; }

 private:
  DeviceBuffer<unsigned char> storage_;
  ParticleQueue view_{};
  std::uint64_t capacity_ = 0;
};

class EmBackend {
 public:
  EmBackend(c8mi_em_config const& config, c8mi_native_table_set const& tables,
            c8mi_environment const& environment, hipStream_t stream);
  ~EmBackend();
  EmBackend(EmBackend const&) = delete;
  EmBackend& operator=(EmBackend const&) = delete;

  void begin_shower(std::uint64_t shower_id, std::uint64_t first_history_id);
  bool can_transport(c8mi_particle_state const& p) const;
  void enqueue(c8mi_particle_state const* particles, std::uint64_t count);
  // species: 0 both (alternating), 1 photons only, 2 leptons only
  void run(int species, std::uint64_t maximum_fronts, c8mi_em_run_result& out);
  void pending(std::uint64_t& photons, std::uint64_t& leptons) const;
  void download_profile(double* photons, double* electrons, double* positrons, double* energy_loss,
                        c8mi_em_ledger* ledger);
  std::uint64_t download_observations(c8mi_particle_state* out, std::uint64_t capacity);
  std::uint64_t download_fallbacks(c8mi_em_fallback* out, std::uint64_t capacity);
  void enable_timing(bool enable);
  void kernel_time(double ms[4], std::uint64_t launches[4]) const;

 private:
  void front(bool photon, c8mi_em_run_result& out);
  void resize_scratch(std::uint64_t n);
  void sync_pending_energy_to(long long* counters);

  c8mi_em_config config_;
  hipStream_t stream_ = nullptr;
  bool own_stream_ = false;
  DeviceTables tables_;
  DeviceBuffer<c8mi_env_node> nodes_;
  DeviceBuffer<std::int32_t> children_;
  DeviceBuffer<c8mi_env_model> models_;
  DeviceBuffer<double> axis_;
  DeviceBuffer<EmContext> context_;
  EmContext host_context_{};
  DeviceBuffer<long long> counters_;
  DeviceBuffer<long long> profile_;
  DeviceBuffer<c8mi_particle_state> observations_;
  DeviceBuffer<c8mi_em_fallback> fallbacks_;
  DeviceBuffer<unsigned long long> record_counts_;
  // queues: [species][buffer]; species 0 photons, 1 leptons
  std::unique_ptr<QueueStorage> queues_[2][2];
  int current_[2] = {0, 0};
  std::uint64_t count_[2] = {0, 0};
  std::uint64_t next_history_ = 0;
  long double injected_ = 0;
  // front scratch
  DeviceBuffer<c8mi_particle_state> staging_;
  DeviceBuffer<std::uint32_t> packed_;
  DeviceBuffer<unsigned char> block_counts_;
  DeviceBuffer<unsigned long long> offsets_;
  DeviceBuffer<FrontTotals> totals_;
  PinnedBuffer<FrontTotals> host_totals_;
  std::uint64_t scratch_capacity_ = 0;
  // timing
  bool timing_ = false;
  std::array<double, 4> ms_{};
  std::array<std::uint64_t, 4> launches_{};
  hipEvent_t events_[4] = {nullptr, nullptr, nullptr, nullptr};
};

}  // namespace c8mi::em
