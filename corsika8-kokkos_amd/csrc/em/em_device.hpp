// Device-side state of the EM backend: the per-shower context every step
// kernel reads (tables, environment, cuts, LPM and thinning parameters, profile
// projection) and the struct-of-arrays particle queues.
//
// The context lives in device memory and is read through a uniform pointer, so
// its fields are fetched by scalar loads shared by the whole wavefront.
#pragma once

#include <hip/hip_runtime.h>

#include "em/hd.hpp"

#include <cstdint>

#include "c8mi/c8mi_em.h"

namespace c8mi::em {

// Ledger / counter layout (int64; energies in profile_energy_scale units per GeV).
enum Counter : int {
  kInjected = 0,
  kContinuous = 1,
  kCutDeposit = 2,
  kCutRestMass = 3,
  kObserved = 4,
  kEscaped = 5,
  kOverlap = 6,
  kMassCorrection = 7,
  kMediumRestMass = 8,
  kPhotoelectricBinding = 9,
  kFallbackEnergy = 10,
  kThinningEnergy = 11,
  kPending = 12,
  kSteps = 13,
  kPhotonCuts = 14,
  kLeptonLimit0 = 15,
  kThinHillas = 23,
  kThinStatistical = 24,
  kThinDiscarded = 25,
  kDepositedSteps = 26,
  kOverflows = 27,
  kInvalid = 28,
  kFallbackReason0 = 29,  // 32 entries
  kNativeMuonCut = 61,    // signed energy of μ-pair muons cut at their first step
  kCounters = 64
};

inline constexpr int kEnergyTerms = 13;
inline constexpr std::uint32_t kMaxProfileBins = 1024;

// Histogram order in the profile buffer.
enum Histogram : int { kPhotonsHist = 0, kElectronsHist = 1, kPositronsHist = 2, kEnergyLossHist = 3 };

// Per-slot shower parameters (one entry per pool slot, device memory).
struct ShowerSlot {
  std::uint64_t key;  // Philox showerId
  std::uint64_t axis_support_count;
  double axis_start[3];
  double axis_direction[3];
  double axis_step;
  double const* axis_grammage;
};

struct EmContext {
  c8mi_native_table_set tables;  // device pointers
  // environment (device pointers)
  c8mi_env_node const* nodes;
  std::int32_t const* children;
  c8mi_env_model const* models;
  std::uint32_t node_count;
  std::uint32_t model_count;
  std::uint32_t observation_enabled;
  std::uint32_t pad0;
  double observation_point[3];
  double observation_normal[3];
  double maximum_magnetic_deflection;
  // physics
  std::uint64_t seed;
  double em_cut_mev;
  double muon_cut_mev;
  c8mi_photon_pair_lpm pair_lpm;
  c8mi_brems_lpm brems_lpm;
  c8mi_em_thinning thinning;
  c8mi_moliere_snapshot moliere;
  // quotients of the Molière integral series, divided once on the host
  // (IEEE division: the values each term of the series would compute)
  double moliere_i1[70];        // c1[i] / (2 i + 1)
  double moliere_i2[70];        // c2[i] / (2 i + 1)
  double moliere_i2_large[13];  // -0.5 c2_large[i] / i
  double moliere_half_over[13]; // 0.5 / i
  std::uint32_t moliere_enabled;
  std::uint32_t moliere_fast_iterations;  // Newton iterations in the step kernel (1..100)
  // interpolation view of the snapshot (null: analytical functions)
  c8mi_moliere_cubic const* moliere_polynomials;  // 4 x 99 cubics
  double const* moliere_guess;                    // 64 x 8 x 128 corrections
  // profile
  std::uint32_t profile_bins;
  std::uint32_t crossing_mode;
  double bin_width;
  double energy_loss_threshold;
  double weight_scale;
  double energy_scale;
  // pool slots and their outputs
  ShowerSlot const* showers;          // shower_slots entries
  std::uint32_t shower_slots;
  std::uint32_t pad1;
  c8mi_particle_state* observations;  // observation_capacity per slot
  std::uint64_t observation_capacity;
  c8mi_em_fallback* fallbacks;        // fallback_capacity per slot
  std::uint64_t fallback_capacity;
  unsigned long long* record_counts;  // per slot: [0] observations, [1] fallbacks
  long long* counters;                // per slot: kCounters
  long long* profile;                 // per slot: 4 * profile_bins
  c8mi_em_first_interaction* first_interactions;  // per slot (null: not captured)
  unsigned long long* first_candidates;           // per slot: final states of generation zero so far
};

// Struct-of-arrays particle queue (ParticleSoA, corsika/kokkos/data/ParticleSoA.hpp:27-82):
// the 13 state fields (116 bytes) plus the pool slot of the particle's shower
// (4 bytes); a queue holds up to a few runs, each sorted by slot.
struct ParticleQueue {
  std::int32_t* pid;
  std::int32_t* medium;
  std::uint32_t* generation;
  std::uint32_t* reserved;
  double* energy;
  double* x;
  double* y;
  double* z;
  double* dx;
  double* dy;
  double* dz;
  double* time;
  double* weight;
  std::uint64_t* history;
  std::uint64_t* parent;
  std::uint64_t* step;
  std::int32_t* node;
  std::uint32_t* shower;
};

C8MI_HD c8mi_particle_state load_particle(ParticleQueue const& q,
                                          std::uint64_t i) {
  c8mi_particle_state p;
  p.pid = q.pid[i];
  p.medium_id = q.medium[i];
  p.generation = q.generation[i];
  p.reserved = q.reserved[i];
  p.energy_gev = q.energy[i];
  p.position_m[0] = q.x[i];
  p.position_m[1] = q.y[i];
  p.position_m[2] = q.z[i];
  p.direction[0] = q.dx[i];
  p.direction[1] = q.dy[i];
  p.direction[2] = q.dz[i];
  p.time_s = q.time[i];
  p.weight = q.weight[i];
  p.history_id = q.history[i];
  p.parent_history_id = q.parent[i];
  p.step_id = q.step[i];
  p.environment_node = q.node[i];
  p.pad_ = 0;
  p.pad2_ = 0;
  return p;
}

C8MI_HD void store_particle(ParticleQueue const& q, std::uint64_t i,
                            c8mi_particle_state const& p) {
  q.pid[i] = p.pid;
  q.medium[i] = p.medium_id;
  q.generation[i] = p.generation;
  q.reserved[i] = p.reserved;
  q.energy[i] = p.energy_gev;
  q.x[i] = p.position_m[0];
  q.y[i] = p.position_m[1];
  q.z[i] = p.position_m[2];
  q.dx[i] = p.direction[0];
  q.dy[i] = p.direction[1];
  q.dz[i] = p.direction[2];
  q.time[i] = p.time_s;
  q.weight[i] = p.weight;
  q.history[i] = p.history_id;
  q.parent[i] = p.parent_history_id;
  q.step[i] = p.step_id;
  q.node[i] = p.environment_node;
}

// Per-source output of a step: up to two particles, each routed to the queue of
// its own species ("next") or of the other species, and either a continuation
// (keeps its history ID) or a new child (history = first + child offset + k).
enum SlotKind : std::uint32_t { kSlotNone = 0, kSlotNextContinue = 1, kSlotNextChild = 2, kSlotOtherChild = 3 };

// Packed per-source counts: bits 0-1 next, 2-3 other, 4-5 children, 6-7 slot0
// kind, 8-9 slot1 kind, 10 a transport step was made, 11 an interaction
// vertex was reached, 12-31 pool slot of the source's shower.
inline constexpr std::uint32_t kMaxShowerSlots = 4096;
inline constexpr int kPackedSlotShift = 12;
__host__ C8MI_HD std::uint32_t pack_counts(std::uint32_t next, std::uint32_t other,
                                           std::uint32_t child, std::uint32_t k0,
                                           std::uint32_t k1, bool step, bool interaction,
                                           std::uint32_t shower) {
  return next | (other << 2) | (child << 4) | (k0 << 6) | (k1 << 8) | (step ? 1u << 10 : 0u) |
         (interaction ? 1u << 11 : 0u) | (shower << kPackedSlotShift);
}

// Front totals written by the scan kernel and read back by the host.
struct FrontTotals {
  unsigned long long next;
  unsigned long long other;
  unsigned long long children;
  unsigned long long steps;
  unsigned long long interactions;
  unsigned long long secondaries;
  unsigned long long pad[2];
};

}  // namespace c8mi::em
