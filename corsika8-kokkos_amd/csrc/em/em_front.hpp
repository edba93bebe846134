// Host launchers of the EM front kernels (em_front.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "em/em_device.hpp"

namespace c8mi::em {

// A queue holds up to kMaxRuns runs, each sorted by pool slot: the previous
// front's stable output, then the secondaries appended by the other species'
// front, then particles staged by the host.
inline constexpr int kMaxRuns = 4;
// Per-(slot, run) output totals of a front: next, other, children, steps,
// interactions.
inline constexpr int kShowerFields = 5;
// Per-slot statistics words: photon steps, lepton steps, photon interactions,
// lepton interactions, secondaries.
inline constexpr int kSlotStats = 8;
// Front report read back by the host: FrontTotals (8 words), then per slot the
// new count of the stepped species and the count appended to the other one.
inline constexpr int kReportHeader = 8;

struct FrontArgs {
  ParticleQueue in;
  std::uint64_t count;
  std::uint64_t run_end[kMaxRuns];  // exclusive ends; runs used: `runs`
  std::uint32_t runs;
  std::uint32_t slots;              // active slot range [0, slots)
  c8mi_particle_state* staging;     // 2 per source
  std::uint32_t* packed;            // 1 per source
  void* block_counts;               // 1 per step block
  unsigned long long* offsets;      // 3 per step block
  unsigned long long* shower_totals;  // [kMaxShowerSlots][kMaxRuns][kShowerFields], zero between fronts
  unsigned long long* adjust;         // [slots][kMaxRuns][3]
  unsigned long long* history;        // [slots] next secondary history ID
  unsigned long long* stats;          // [slots][kSlotStats]
  unsigned long long* report;         // kReportHeader + 2 * slots
};

std::size_t step_lds_bytes(std::uint32_t bins);
std::size_t block_count_bytes(std::uint64_t n);
std::size_t offset_bytes(std::uint64_t n);

void launch_step(bool photon, EmContext const* context, std::uint32_t bins, FrontArgs const& f,
                 hipStream_t stream);
void launch_scan(bool photon, FrontArgs const& f, hipStream_t stream);
void launch_scatter(FrontArgs const& f, ParticleQueue const& next_queue, ParticleQueue const& other_queue,
                    std::uint64_t other_base, hipStream_t stream);

}  // namespace c8mi::em
