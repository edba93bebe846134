// Host side of the EM backend: device storage of tables, environment and
// context; per-species double-buffered particle queues shared by a pool of
// shower slots; the front loop.
// Mirrors the reference's IEmBackend implementation
// (corsika/kokkos/em/execution/IEmBackend.hpp:25-88,
// corsika/detail/kokkos/em/execution/BackendInstanceImplementation.inl) with
// one HIP stream per backend and one small device-to-host read per front. The
// IEmBackend-shaped calls (begin_shower, enqueue, run, downloads) act on slot 0
// of the pool; the slot_* calls keep several showers in the same fronts.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <memory>
#include <vector>

#include "c8mi/c8mi_em.h"
#include "em/device_tables.hpp"
#include "em/em_device.hpp"
#include "em/em_front.hpp"
#include "em/em_order.hpp"
#include "radio/radio_accumulator.hpp"
#include "radio/slot_radio.hpp"
#include "runtime.hpp"

namespace c8mi::em {

// Checks of the backend's construction arguments (std::invalid_argument).
void validate_environment(c8mi_environment const& environment);
void validate_config(c8mi_em_config const& config);
// The context's scalar fields from the configuration and environment (every
// pointer field is left to the caller).
void fill_context(c8mi_em_config const& config, c8mi_environment const& environment, EmContext& context);
// The ledger of a shower's int64 counters (energies divided by the scale).
void ledger_from_counters(long long const* counters, double energy_scale, c8mi_em_ledger& out);

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

  // IEmBackend-shaped single-shower calls: the whole pool is reset and the
  // shower runs in slot 0.
  void begin_shower(std::uint64_t shower_id, std::uint64_t first_history_id);
  bool can_transport(c8mi_particle_state const& p) const;
  void enqueue(c8mi_particle_state const* particles, std::uint64_t count) { slot_enqueue(0, particles, count); }
  // Rounds of the reference's call schedule over every active slot (each
  // shower: a photon call of at most 16 fronts, then a lepton call of at most
  // 1024, with history ranges reserved from its counter), until nothing is
  // pending or `maximum_rounds` rounds (a photon front, then a lepton front)
  // have run.
  void run(std::uint64_t maximum_rounds, c8mi_em_run_result& out);
  // IEmBackend::runPhotonWavefront / runLeptonWavefront for the shower of slot
  // 0 (run_photon_cascade, PhotonCascade.inl:177-574; run_lepton_cascade,
  // LeptonCascade.inl:86-565): one call over the device pending prefix and
  // `particles`; the particles of the species left when the call ends are
  // downloaded and removed (remaining()).
  void run_call(bool photon, c8mi_particle_state const* particles, std::uint64_t count,
                std::uint64_t first_secondary_history_id, std::uint64_t maximum_wavefronts,
                std::uint64_t history_limit_exclusive, std::uint64_t minimum_resident_batch,
                c8mi_em_wavefront_result& out);
  std::vector<c8mi_particle_state> const& remaining() const noexcept { return remaining_; }
  void limits(c8mi_em_limits& out) const;
  void statistics(c8mi_em_statistics& out) const;
  void pending(std::uint64_t& photons, std::uint64_t& leptons) const;
  void download_profile(double* photons, double* electrons, double* positrons, double* energy_loss,
                        c8mi_em_ledger* ledger) {
    slot_download_profile(0, photons, electrons, positrons, energy_loss, ledger);
  }
  std::uint64_t download_observations(c8mi_particle_state* out, std::uint64_t capacity) {
    return slot_download_records(0, 0, out, sizeof(c8mi_particle_state), capacity);
  }
  std::uint64_t download_fallbacks(c8mi_em_fallback* out, std::uint64_t capacity) {
    return slot_download_records(0, 1, out, sizeof(c8mi_em_fallback), capacity);
  }

  // pool
  std::uint32_t slots() const noexcept { return slots_; }
  void slot_begin(std::uint32_t slot, std::uint64_t shower_id, std::uint64_t first_history_id,
                  c8mi_em_shower_axis const* axis);
  void slot_enqueue(std::uint32_t slot, c8mi_particle_state const* particles, std::uint64_t count);
  void slot_state(std::uint32_t slot, c8mi_em_slot_state& out) const;
  void slot_download_profile(std::uint32_t slot, double* photons, double* electrons, double* positrons,
                             double* energy_loss, c8mi_em_ledger* ledger);
  std::uint64_t slot_download_records(std::uint32_t slot, int kind, void* out, std::size_t record_bytes,
                                      std::uint64_t capacity);
  // The slot's first-interaction snapshot; false when none was captured.
  bool slot_download_first_interaction(std::uint32_t slot, c8mi_em_first_interaction& out);
  void slot_end(std::uint32_t slot);
  void slot_pending(std::uint32_t slot, std::uint64_t& photons, std::uint64_t& leptons) const;
  void serve(c8mi_em_request const* requests, std::uint64_t count, c8mi_em_serve_options const& options,
             c8mi_em_request_result* results, c8mi_em_run_result& totals);

  // Radio (IEmBackend::gpuRadioEnabled / downloadRadioWaveforms with the
  // backend's RadioConfig): every e+- step of every lepton front becomes a
  // compact track projected onto `radio` (null detaches). The projection runs on
  // the accumulator's own stream, overlapping the following fronts; two track
  // buffers alternate between fronts. The accumulator must outlive the
  // attachment.
  void attach_radio(radio::RadioAccumulator* radio);

  // Per-shower radio of the pool (IEmBackend's radio per shower, here for
  // every slot at once): each slot's e+- steps are projected onto that slot's
  // own observers, waveforms and counters, resident on the device.
  void enable_slot_radio(c8mi_radio_config const& config, std::uint32_t observers_per_slot,
                         std::uint32_t bins_per_observer);
  void slot_set_radio(std::uint32_t slot, c8mi_radio_observer const* observers, std::uint64_t count);
  void slot_download_radio(std::uint32_t slot, std::int64_t* coreas, std::int64_t* zhs,
                           c8mi_radio_counters* counters);
  void slot_radio_time(double& ms, std::uint64_t& launches);

  void enable_timing(bool enable);
  void kernel_time(double ms[kFrontKernels], std::uint64_t launches[kFrontKernels]) const;

 private:
  using Staged = StagedParticle;
  // A shower's place in the reference's call schedule (EmRouter.inl:339-416).
  struct CallState {
    int phase = 0;                  // species of the current or next call
    bool open = false;              // a call of `phase` holds a history range
    bool opening = false;           // its first front has not run yet
    int stop = 0;                   // how the last call ended (kStop*)
    std::uint64_t call_fronts = 0;  // fronts of the open call
    std::uint64_t counter = 0;      // next history ID to reserve
    std::uint64_t next = 0;         // next secondary history ID of the open call
    std::uint64_t limit = 0;        // exclusive limit of the open call
    std::uint64_t cap = 0;          // fronts per call
    std::uint64_t min_batch = 0;    // below-minimum-batch checkpoint
    bool resident = true;           // calls open by themselves (false during a wavefront call)
  };
  struct SlotHost {
    bool active = false;
    std::uint64_t shower_id = 0;
    std::uint64_t pending[2] = {0, 0};  // on the device
    std::uint64_t staged[2] = {0, 0};   // on the host
    std::uint64_t fronts[2] = {0, 0};
    long double injected = 0;
    DeviceBuffer<double> axis;          // own projection table, if any
    CallState call;
  };
  enum : int { kStopEmpty = 0, kStopCap = 1, kStopMinimumBatch = 2, kStopHistory = 3 };

  void reset_pool();
  PoolArrays pool_arrays() const;
  // Enqueues the start of the given slots (no host synchronisation).
  void start_slots(SlotInit const* inits, std::uint32_t count);
  void activate(std::uint32_t slot, std::uint64_t shower_id, std::uint64_t first_history_id);
  void decode_ledger(long long* counters, std::uint32_t slot, c8mi_em_ledger& out) const;
  // One front of one species over the slots whose call runs now; false when
  // no slot's call ran a front.
  bool front(bool photon, c8mi_em_run_result& out);
  // Whether slot `h`'s call of `species` runs a front over its n particles
  // now (opening or ending calls as the schedule says).
  bool settle(CallState& c, std::uint64_t const n[2], c8mi_em_run_result& out) const;
  void reserve_call(CallState& c, int species, std::uint64_t n, c8mi_em_run_result& out) const;
  void run_rounds_loop(std::uint64_t maximum_rounds, c8mi_em_run_result& out);
  void flush_staged(int species);
  void add_medium(std::int32_t medium);
  std::uint32_t active_slot_count() const;
  void resize_scratch(std::uint64_t n);
  void check_slot(std::uint32_t slot) const;
  void update_active_range();
  long long pending_energy(std::uint32_t slot);
  void wait_radio_tracks();
  void create_track_events();

  c8mi_em_config config_;
  std::uint32_t slots_ = 1;
  std::uint32_t active_range_ = 0;  // 1 + highest active slot
  hipStream_t stream_ = nullptr;
  bool own_stream_ = false;
  DeviceTables tables_;
  DeviceBuffer<c8mi_env_node> nodes_;
  DeviceBuffer<std::int32_t> children_;
  DeviceBuffer<c8mi_env_model> models_;
  DeviceBuffer<double> axis_;
  DeviceBuffer<c8mi_moliere_cubic> moliere_polynomials_;
  DeviceBuffer<double> moliere_guess_;
  DeviceBuffer<EmContext> context_;
  EmContext host_context_{};
  DeviceBuffer<ShowerSlot> showers_;
  std::vector<ShowerSlot> host_showers_;
  std::vector<SlotHost> slot_host_;
  DeviceBuffer<long long> counters_;
  DeviceBuffer<long long> profile_;
  DeviceBuffer<c8mi_particle_state> observations_;
  DeviceBuffer<c8mi_em_fallback> fallbacks_;
  DeviceBuffer<unsigned long long> record_counts_;
  DeviceBuffer<c8mi_em_first_interaction> first_interactions_;
  DeviceBuffer<unsigned long long> first_candidates_;
  DeviceBuffer<unsigned long long> history_;
  DeviceBuffer<unsigned long long> stats_;
  DeviceBuffer<unsigned long long> shower_totals_;
  DeviceBuffer<unsigned long long> adjust_;
  DeviceBuffer<unsigned long long> report_;
  PinnedBuffer<unsigned long long> host_report_;
  // batched slot starts and result gathers
  DeviceBuffer<SlotInit> init_dev_;
  PinnedBuffer<SlotInit> init_host_;
  DeviceBuffer<std::uint32_t> gather_slots_dev_;
  PinnedBuffer<std::uint32_t> gather_slots_host_;
  DeviceBuffer<unsigned long long> gather_dev_;
  PinnedBuffer<unsigned long long> gather_host_;
  hipEvent_t init_copied_ = nullptr;
  DeviceBuffer<StagedParticle> unstage_dev_;
  PinnedBuffer<StagedParticle> unstage_host_;
  hipEvent_t unstage_copied_ = nullptr;
  // queues: [species][buffer]; species 0 photons, 1 leptons
  std::unique_ptr<QueueStorage> queues_[2][2];
  int current_[2] = {0, 0};
  std::uint64_t count_[2] = {0, 0};
  // the queue holds only the last front's continuing particles (clean slot
  // words, in front order): a front needs no ordering step unless a call
  // opens, a shower is bucketed or sits out
  bool clean_[2] = {true, true};
  std::vector<Staged> staged_[2];
  // front ordering (bucketing, call-input order, parked showers)
  FrontOrder order_;
  std::vector<std::int32_t> media_;  // sorted distinct medium IDs
  DeviceBuffer<SlotControl> control_dev_;
  PinnedBuffer<SlotControl> control_host_;
  hipEvent_t control_copied_ = nullptr;
  std::vector<c8mi_particle_state> remaining_;
  std::vector<CallState> plan_;
  std::vector<std::uint64_t> front_counts_;
  // cumulative statistics (IEmBackend::statistics)
  c8mi_em_statistics totals_{};
  // front scratch
  DeviceBuffer<c8mi_particle_state> staging_;
  DeviceBuffer<std::uint32_t> packed_;
  DeviceBuffer<unsigned char> block_counts_;
  DeviceBuffer<unsigned long long> offsets_;
  DeviceBuffer<VertexRecord> vertices_;
  DeviceBuffer<unsigned char> moves_;
  DeviceBuffer<std::uint32_t> moved_;
  DeviceBuffer<unsigned char> moliere_open_;
  DeviceBuffer<unsigned long long> moliere_open_count_;
  DeviceBuffer<VertexRecord> open_vertices_;
  DeviceBuffer<unsigned long long> open_vertex_count_;
  // the open Molière samples' stream (tail, their finish and vertices)
  hipStream_t side_stream_ = nullptr;
  hipEvent_t open_ready_ = nullptr;
  hipEvent_t open_done_ = nullptr;
  DeviceBuffer<unsigned long long> vertex_count_;
  std::uint64_t scratch_capacity_ = 0;
  // radio: tracks of the last two lepton fronts; `track_done_` marks the end
  // of their projection on the accumulator's stream
  radio::RadioAccumulator* radio_ = nullptr;
  std::unique_ptr<radio::SlotRadio> slot_radio_;
  DeviceBuffer<double> tracks_[2];
  std::uint64_t track_capacity_[2] = {0, 0};
  hipEvent_t track_ready_[2] = {nullptr, nullptr};
  hipEvent_t track_done_[2] = {nullptr, nullptr};  // attached accumulator
  hipEvent_t slot_done_[2] = {nullptr, nullptr};   // per-shower radio
  bool track_busy_[2] = {false, false};
  bool slot_busy_[2] = {false, false};
  int track_parity_ = 0;
  // per-shower radio: (slot, run) key arrays and the segment list of a front
  DeviceBuffer<unsigned long long> segment_keys_[2];  // begin [slots * kMaxRuns], then end
  DeviceBuffer<radio::SegmentDev> segments_[2];
  DeviceBuffer<std::uint32_t> segment_count_[2];
  DeviceBuffer<unsigned long long> item_count_[2];
  // timing
  bool timing_ = false;
  std::array<double, kFrontKernels> ms_{};
  std::array<std::uint64_t, kFrontKernels> launches_{};
  hipEvent_t events_[kFrontKernels + 5] = {};  // up to 9 marks per front, and the one before them
};

}  // namespace c8mi::em
