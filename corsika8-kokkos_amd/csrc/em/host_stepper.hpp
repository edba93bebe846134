// The host's scalar stepper: one step of one EM particle on the CPU, the role
// the reference's ScalarCascadeStepper (corsika/kokkos/scheduling/
// ScalarCascadeStepper.hpp) plays for the particles EmRouter returns to the
// host (EmRouter.inl:260-293). It runs the device path's own step functions
// (em/steps.hpp, compiled for the host) on host copies of the tables and the
// environment, so a particle stepped here takes the step a front would give
// it: the same Philox keys, decisions and outputs. Profile, ledger and records
// go to the stepper's own accumulators (one shower at a time, one thread).
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "c8mi/c8mi_em.h"
#include "em/device_tables.hpp"
#include "em/em_device.hpp"

namespace c8mi::em {

struct HostStep {
  c8mi_particle_state out[2];  // emission order: a continuation alone, or children by history ID
  std::uint32_t count;
  std::uint32_t children;      // new history IDs taken (first_history_id + k)
  bool step;                   // a transport record was made
  bool interaction;            // the step reached an interaction vertex
  bool has_track;              // an e+- transport step: `track` holds its compact radio track
  double track[9];
};

class HostStepper {
 public:
  HostStepper(c8mi_em_config const& config, c8mi_native_table_set const& tables,
              c8mi_environment const& environment);
  // Shares a host copy of the tables (read only) with other steppers.
  HostStepper(c8mi_em_config const& config, std::shared_ptr<HostTables const> tables,
              c8mi_environment const& environment);
  HostStepper(HostStepper const&) = delete;
  HostStepper& operator=(HostStepper const&) = delete;

  // Starts a shower (Philox showerId) and zeroes the accumulators.
  void begin_shower(std::uint64_t shower_id);
  // moliere false: the step without Molière scattering (the host's step of
  // a particle whose device step failed the Molière sampling).
  HostStep step(c8mi_particle_state const& particle, std::uint64_t first_history_id, bool moliere = true);
  // EmRouter::canRoute for EM particles (EmRouter.inl:64-105): an e+-/photon
  // the device can transport (EmBackend::can_transport) inside the environment.
  bool can_route(c8mi_particle_state const& particle) const;
  // Adds energy * weight to the ledger's injected term (negative: energy the
  // device took over), rounded like every ledger term.
  void book_injected(double energy_weight);

  std::uint32_t bins() const noexcept { return context_.profile_bins; }
  long long const* counters() const noexcept { return counters_.data(); }
  long long const* profile() const noexcept { return profile_.data(); }
  std::uint64_t observation_count() const noexcept { return record_counts_[0]; }
  std::uint64_t fallback_count() const noexcept { return record_counts_[1]; }
  c8mi_particle_state const* observations() const noexcept { return observations_.data(); }
  c8mi_em_fallback const* fallbacks() const noexcept { return fallbacks_.data(); }
  bool first_interaction(c8mi_em_first_interaction& out) const {
    if (first_candidates_ == 0) return false;
    out = first_interaction_;
    return true;
  }
  c8mi_em_config const& config() const noexcept { return config_; }

 private:
  static constexpr std::uint64_t kInitialRecords = 64;
  void reserve_records();

  c8mi_em_config config_;
  std::shared_ptr<HostTables const> tables_;
  std::vector<c8mi_env_node> nodes_;
  std::vector<std::int32_t> children_;
  std::vector<c8mi_env_model> models_;
  std::vector<double> axis_;
  std::vector<c8mi_moliere_cubic> polynomials_;
  std::vector<double> guess_;
  std::vector<long long> counters_;
  std::vector<long long> profile_;
  std::vector<c8mi_particle_state> observations_;
  std::vector<c8mi_em_fallback> fallbacks_;
  unsigned long long record_counts_[2] = {0, 0};
  c8mi_em_first_interaction first_interaction_{};
  unsigned long long first_candidates_ = 0;
  ShowerSlot shower_{};
  EmContext context_{};
};

}  // namespace c8mi::em
