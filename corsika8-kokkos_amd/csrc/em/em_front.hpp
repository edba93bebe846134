// Host launchers of the EM front kernels (em_front.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "em/em_device.hpp"
#include "radio/radio_device.hpp"

namespace c8mi::em {

// A front's input is one run sorted by pool slot (the previous front's stable
// output, or the gather of the ordering step); the (slot, run) arrays keep
// room for kMaxRuns runs.
inline constexpr int kMaxRuns = 4;
// Per-(slot, run) output totals of a front: next, other, children, steps,
// interactions.
inline constexpr int kShowerFields = 5;
// Per-slot statistics words: photon steps, lepton steps, photon interactions,
// lepton interactions, secondaries.
inline constexpr int kSlotStats = 8;
// Front report read back by the host: FrontTotals (8 words), then per slot the
// new count of the stepped species, the count appended to the other one and
// the children given history IDs.
inline constexpr int kReportHeader = 8;
inline constexpr int kReportPerSlot = 3;

// An e+- step that reached its interaction vertex, handed from the transport
// kernel to the vertex kernel (compacted, in source order within each block).
struct VertexRecord {
  c8mi_particle_state parent;
  std::uint64_t shower_key;
  std::uint64_t source;
  double total_rate;
  double process_uniform;
  double particle_mass_gev;
  double vertex_density;
  std::uint32_t slot;
  std::uint32_t pad[3];
};

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
  VertexRecord* vertices;             // lepton fronts: up to `count` records
  unsigned long long* vertex_count;   // zero between fronts
  void* moves;                        // split lepton fronts: one LeptonMove per source
  std::uint32_t* moved;               // 1 where the move kernel produced a move
  double* tracks;                     // lepton fronts with radio: [9][track_stride] compact tracks
  std::uint64_t track_stride;
  void* moliere_open;                 // split lepton fronts: up to `count` open Molière samples
  unsigned long long* moliere_open_count;  // zeroed by the move kernel
  VertexRecord* open_vertices;        // vertices of the sources finished after the tail kernel
  unsigned long long* open_vertex_count;  // zero between fronts
  c8mi_em_first_interaction* first_interactions;  // per slot; the scatter gives a photoelectric
  unsigned long long const* first_candidates;     // snapshot's child its assigned history ID
};

// States of a source between the split lepton kernels (FrontArgs::moved).
inline constexpr std::uint32_t kMovedNone = 0;           // no move: the move kernel recorded a fallback
inline constexpr std::uint32_t kMovedOk = 1;             // moved (and scattered, when Molière is on)
inline constexpr std::uint32_t kMovedMoliereFailed = 2;  // the scattering failed: a transport fallback
inline constexpr std::uint32_t kMovedOpen = 3;           // the sample continues in the tail kernel

// Start of a shower in a slot: its parameters and first secondary history ID.
struct SlotInit {
  ShowerSlot shower;
  unsigned long long first_history;
  std::uint32_t slot;
  std::uint32_t pad;
};

// Words gathered per finished slot: ledger counters, statistics, next history
// ID, record counts.
inline constexpr int kGatherCounters = 0;
inline constexpr int kGatherStats = kCounters;
inline constexpr int kGatherHistory = kCounters + kSlotStats;
inline constexpr int kGatherRecords = kGatherHistory + 1;
inline constexpr int kGatherWords = 80;

struct PoolArrays {
  ShowerSlot* showers;
  long long* profile;
  long long* counters;
  unsigned long long* record_counts;
  unsigned long long* first_candidates;
  unsigned long long* stats;
  unsigned long long* history;
  std::uint32_t bins;
};

std::size_t step_lds_bytes(std::uint32_t bins);
std::size_t block_count_bytes(std::uint64_t n);
std::size_t offset_bytes(std::uint64_t n);

// Kernel time categories of a front.
// The split e+- form books its move and finish kernels as the transport.
enum FrontKernel : int { kKernPhotonStep = 0, kKernLeptonTransport = 1, kKernLeptonVertex = 2, kKernCount = 3,
                         kKernScan = 4, kKernScatter = 5, kFrontKernels = 6 };

void launch_photon_step(EmContext const* context, std::uint32_t bins, FrontArgs const& f, hipStream_t stream);
void launch_lepton_transport(EmContext const* context, std::uint32_t bins, FrontArgs const& f, hipStream_t stream);
// Split form of the e+- transport (when Molière scattering is on): the move,
// then the finish (scattering, profile, outputs, vertex compaction).
std::size_t lepton_move_bytes();
std::size_t moliere_open_bytes();
void launch_lepton_move(EmContext const* context, std::uint32_t bins, FrontArgs const& f, hipStream_t stream);
// Per-shower radio: the segment list of a lepton front's tracks for
// radio::SlotRadio::project (key arrays [slots][kMaxRuns], zero between uses).
void launch_radio_segments(FrontArgs const& f, std::uint32_t slots, unsigned long long* begin,
                           unsigned long long* end, std::uint32_t const* observer_counts, std::uint64_t chunk,
                           std::uint32_t tiles, radio::SegmentDev* segments, std::uint32_t* segment_count,
                           unsigned long long* item_count, hipStream_t stream);

// Molière scattering of the moved sources between the move and finish kernels
// (the step kernel, then the tail kernel for samples still open);
// table: the interpolation-view sampler (else the analytical functions).
void launch_lepton_moliere(EmContext const* context, bool table, FrontArgs const& f, hipStream_t stream);
// The open Molière samples on a stream of their own, overlapping the finish
// and vertex kernels of the other sources: tail kernel, their finish, their
// vertices (FrontArgs::open_vertices).
void launch_lepton_open(EmContext const* context, std::uint32_t bins, bool table, FrontArgs const& f,
                        hipStream_t stream);
void launch_lepton_finish(EmContext const* context, std::uint32_t bins, FrontArgs const& f, hipStream_t stream);
void launch_lepton_vertex(EmContext const* context, std::uint32_t bins, FrontArgs const& f, hipStream_t stream);
void launch_count(FrontArgs const& f, hipStream_t stream);
void launch_scan(bool photon, FrontArgs const& f, hipStream_t stream);
void launch_scatter(FrontArgs const& f, ParticleQueue const& next_queue, ParticleQueue const& other_queue,
                    std::uint64_t other_base, hipStream_t stream);
// Host-staged particle with the pool slot of its shower.
struct StagedParticle {
  c8mi_particle_state particle;
  std::uint32_t slot;
  std::uint32_t pad[3];
};

// Copies queue entries [0, n) into `out` as particle records.
void launch_pack(ParticleQueue const& queue, std::uint64_t n, c8mi_particle_state* out, hipStream_t stream);

// Writes n staged particles into queue entries [at, at + n), class host.
void launch_unstage(StagedParticle const* staged, std::uint64_t n, ParticleQueue const& queue, std::uint64_t at,
                    hipStream_t stream);
// Clears the profile, ledger, records and statistics of each started slot and
// sets its parameters and history counter (one workgroup per slot).
void launch_slot_init(PoolArrays const& pool, SlotInit const* inits, std::uint32_t count, hipStream_t stream);
// Copies kGatherWords per listed slot into `out` (one workgroup per slot).
void launch_slot_gather(PoolArrays const& pool, std::uint32_t const* slots, std::uint32_t count,
                        unsigned long long* out, hipStream_t stream);

}  // namespace c8mi::em
