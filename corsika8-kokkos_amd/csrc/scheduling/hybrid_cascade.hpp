// Host handoff of an EM shower: the reference's HybridCascade::run
// (corsika/detail/kokkos/scheduling/HybridCascade.inl:55-175) with EmRouter
// (corsika/detail/kokkos/scheduling/EmRouter.inl:64-423) for EM particles,
// over the EM C ABI. A LIFO host stack (CpuParticleScheduler::acquireNext,
// CpuParticleScheduler.inl:27-52) feeds the router: a particle it can route
// is staged; any other, and each particle the router returns for a host step,
// takes one step on the host stepper and its outputs go back on the stack.
// The router hands the staged batch to the device when it reaches the
// minimum batch, when the device holds pending particles, or after the
// maximum run of consecutive expansion rounds; a smaller batch is returned to
// the stack for one host step each (scalar expansion, EmRouter.inl:260-293).
// Device calls are the IEmBackend wavefront calls with history ranges
// reserved from the shower's counter (EmRouter.inl:336-416,640-679); their
// remaining particles are staged again.
#pragma once

#include <cstdint>
#include <set>
#include <unordered_map>
#include <utility>
#include <vector>

#include "c8mi/c8mi_hybrid.h"
#include "em/host_stepper.hpp"

namespace c8mi::scheduling {

// The reference's constants for every zero field (or for a null config).
c8mi_em_hybrid_config resolve_config(c8mi_em_hybrid_config const* config);

class HybridCascade {
 public:
  // em may be null for prefix-only use.
  HybridCascade(c8mi_em* em, em::HostStepper& stepper, c8mi_em_hybrid_config const* config);

  void attach_radio(c8mi_radio* radio) { radio_ = radio; }
  c8mi_em_hybrid_result run(c8mi_particle_state const* primaries, std::uint64_t count, std::uint64_t shower_id,
                            std::uint64_t first_history_id);
  // The shower up to its first device call: the batch the router would hand
  // over, the compact tracks of the host e+- steps ([n][9]) and the counter.
  c8mi_em_hybrid_result prefix(c8mi_particle_state const* primaries, std::uint64_t count, std::uint64_t shower_id,
                               std::uint64_t first_history_id, std::vector<c8mi_particle_state>& batch,
                               std::vector<double>& tracks, std::uint64_t& next_history_id);

 private:
  void start(c8mi_particle_state const* primaries, std::uint64_t count, std::uint64_t shower_id,
             std::uint64_t first_history_id);
  void cascade();
  enum class Route { kDevice, kHostStep, kFallbackStep };
  Route route(c8mi_particle_state const& p);
  void return_fallbacks();
  void stage(c8mi_particle_state const& p);
  void host_step(c8mi_particle_state const& p, bool moliere);
  std::uint64_t device_pending() const;
  bool ready() const;
  bool pending() const;
  void advance();
  void device_call(bool lepton, std::vector<c8mi_particle_state> const& host);
  void flush_tracks();

  c8mi_em* em_;
  em::HostStepper& stepper_;
  c8mi_em_hybrid_config config_;
  c8mi_em_limits limits_{};
  c8mi_radio* radio_ = nullptr;
  std::vector<c8mi_particle_state> stack_;
  std::vector<c8mi_particle_state> staged_;
  std::vector<c8mi_particle_state> remaining_;
  // cpuWavefrontExpansionSteps_: (history, step) of the particles returned
  // for a host step; canRoute declines each once.
  std::set<std::pair<std::uint64_t, std::uint64_t>> expansion_steps_;
  // cpuFallbackSteps_: device fallbacks returned for a host step, which runs
  // without Molière scattering.
  std::set<std::pair<std::uint64_t, std::uint64_t>> fallback_steps_;
  std::vector<c8mi_em_fallback> fallbacks_;
  std::uint64_t fallback_cursor_ = 0;
  // lastCpuExpansionStates_: a staged particle in the state of its last
  // expansion forces the next small batch to the device.
  std::unordered_map<std::uint64_t, c8mi_particle_state> expanded_;
  std::vector<double> tracks_;  // host e+- tracks not yet handed to the radio, [n][9]
  std::vector<c8mi_lepton_transport_record> records_;
  bool force_small_ = false;
  std::uint64_t consecutive_ = 0;
  std::uint64_t counter_ = 0;
  bool prefix_only_ = false;
  bool stop_ = false;
  double host_seconds_ = 0., device_seconds_ = 0.;
  c8mi_em_hybrid_result r_{};
};

}  // namespace c8mi::scheduling
