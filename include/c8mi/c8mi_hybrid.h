/*
 * c8mi host handoff C ABI — the CPU side of an EM shower that enters the
 * device path through the reference's router: the scalar stepper that runs
 * single particle steps on the host, and the cascade driver that stages
 * particles, expands small batches on the host and hands batches to the
 * device through the EM C ABI (c8mi_em.h), as the reference's HybridCascade
 * (corsika/detail/kokkos/scheduling/HybridCascade.inl:55-175) drives EmRouter
 * (corsika/detail/kokkos/scheduling/EmRouter.inl:64-423) for electromagnetic
 * particles.
 *
 * The host steps run the device path's own step functions compiled for the
 * host (same tables, same Philox keys), so a particle stepped on the host
 * takes the step the device would give it. Profile and ledger contributions
 * of host steps go to the stepper's own accumulators (the reference's C8 CPU
 * output), those of device steps to the EM handle's.
 *
 * Every entry point returns a c8mi_status (0 = success); on failure
 * c8mi_last_error() holds the message for the calling thread.
 */
#ifndef C8MI_HYBRID_H
#define C8MI_HYBRID_H

#include "c8mi_em.h"
#include "c8mi_radio.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scalar stepper (ScalarCascadeStepper::advance, one CPU step) ---- */

typedef struct c8mi_em_scalar c8mi_em_scalar;

/* The arguments of c8mi_em_create (deep-copied; the stepper holds no device
 * memory and runs without a device). */
int c8mi_em_scalar_create(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                          const c8mi_environment* environment, c8mi_em_scalar** out);
int c8mi_em_scalar_destroy(c8mi_em_scalar* scalar);

/* Starts a shower (its Philox showerId) and zeroes the accumulators. */
int c8mi_em_scalar_begin_shower(c8mi_em_scalar* scalar, uint64_t shower_id);

/* One step of one particle (photon: PhotonFront.inl:137-397 for one source;
 * e+-: LeptonFront.inl:226-550 with Molière scattering iterated to
 * convergence): out[0..*count) are its outputs in emission order, either one
 * continuation (same history ID, next step ID) or *children new particles
 * with history IDs first_history_id + k. When the step is an e+- transport
 * step and `track` is not NULL, its compact radio track (c8mi_types.h order:
 * start xyz, end xyz, start and end time, charge-signed weight) is written
 * and *has_track set. A step the device would hand to the host (a fallback)
 * is recorded as a fallback record and has no outputs. */
int c8mi_em_scalar_step(c8mi_em_scalar* scalar, const c8mi_particle_state* particle,
                        uint64_t first_history_id, c8mi_particle_state out[2], uint32_t* count,
                        uint32_t* children, double track[9], int* has_track);

/* The host accumulators since the last begin: histograms as
 * c8mi_em_download_profile (NULL skips one), the ledger, and the records. */
int c8mi_em_scalar_download_profile(const c8mi_em_scalar* scalar, double* photons, double* electrons,
                                    double* positrons, double* energy_loss, c8mi_em_ledger* ledger);
int c8mi_em_scalar_download_observations(const c8mi_em_scalar* scalar, c8mi_particle_state* out,
                                         uint64_t capacity, uint64_t* count);
int c8mi_em_scalar_download_fallbacks(const c8mi_em_scalar* scalar, c8mi_em_fallback* out,
                                      uint64_t capacity, uint64_t* count);
/* The shower's first interaction when its generation-zero particle reached a
 * final state in a host step (as c8mi_em_download_first_interaction). */
int c8mi_em_scalar_download_first_interaction(const c8mi_em_scalar* scalar, c8mi_em_first_interaction* out,
                                              int* captured);

/* ---- cascade driver (HybridCascade + EmRouter for EM particles) ---- */

/* EmRouter's scheduling constants. Zero fields take the reference's values. */
typedef struct c8mi_em_hybrid_config {
  uint64_t minimum_batch_size;            /* staging threshold of a device call and the calls'
                                             minimum resident batch (BackendConfig.hpp:26, 16384) */
  uint32_t maximum_expansion_rounds;      /* consecutive host expansions before a small batch is
                                             forced to the device (EmRouter.inl:1262, 8) */
  uint32_t photon_wavefronts;             /* per photon call (EmRouter.inl:368, 16) */
  uint32_t lepton_wavefronts;             /* per lepton call (EmRouter.inl:405, 1024) */
  uint32_t lepton_reservation_multiplier; /* EmRouter.inl:670, 64 */
  uint64_t lepton_reservation_chunk;      /* EmRouter.inl:671, 2^20 */
} c8mi_em_hybrid_config;

/* One shower's counts (EmRouterStatistics, HybridCascade timing). */
typedef struct c8mi_em_hybrid_result {
  /* host steps */
  uint64_t cpu_particles;          /* particles stepped on the host */
  uint64_t cpu_photon_steps;       /* transport records of host steps */
  uint64_t cpu_lepton_steps;
  uint64_t cpu_interactions;
  uint64_t cpu_secondaries;        /* children of host steps */
  uint64_t cpu_tracks;             /* e+- tracks of host steps (handed to the radio) */
  /* router */
  uint64_t particles_staged;       /* stage() calls */
  uint64_t small_batch_expansions; /* host expansion rounds */
  uint64_t particles_returned_for_expansion;
  uint64_t scalar_expansion_budget_flushes; /* small batches forced to the device */
  uint64_t stalled_boundary_flushes;        /* forced by an unchanged expanded state */
  uint64_t wavefronts;             /* advances that called the device */
  uint64_t input_batch_splits;
  uint64_t maximum_input_batch;
  uint64_t reserved_history_ids;
  uint64_t below_minimum_batch_checkpoints;
  uint64_t history_range_checkpoints;
  /* device calls */
  uint64_t photon_calls;
  uint64_t lepton_calls;
  uint64_t device_photon_fronts;
  uint64_t device_lepton_fronts;
  uint64_t device_photon_steps;
  uint64_t device_lepton_steps;
  uint64_t device_interactions;
  uint64_t device_secondaries;
  uint64_t bucketed_fronts;
  uint64_t next_history_id;        /* the shower's history counter at the end */
  /* fallback round trip (EmRouter.inl:709-800): device fallback records whose
   * particle went back to the host stack for a host step (Molière sampling
   * failures: the step again without Molière scattering), and those kept as
   * records (specified final states the host has no model for) */
  uint64_t fallbacks_returned;
  uint64_t fallbacks_kept;
  double host_seconds;             /* wall time of host steps */
  double device_seconds;           /* wall time of device calls (remaining downloads included) */
} c8mi_em_hybrid_result;

typedef struct c8mi_em_hybrid c8mi_em_hybrid;

/* A driver over an EM handle and a scalar stepper built from the same
 * configuration, tables and environment (neither is owned; both must outlive
 * the driver). NULL config: the reference's constants. */
int c8mi_em_hybrid_create(c8mi_em* em, c8mi_em_scalar* scalar, const c8mi_em_hybrid_config* config,
                          c8mi_em_hybrid** out);
int c8mi_em_hybrid_destroy(c8mi_em_hybrid* hybrid);

/* Host radio tracks (RadioAccumulator::enqueueHostLeptonTracks,
 * RadioAccumulator.inl:252-290): the e+- tracks of host steps are enqueued on
 * `radio` after each expansion round; NULL detaches. Attach the same handle to
 * the EM handle (c8mi_em_attach_radio) for the device steps. */
int c8mi_em_hybrid_attach_radio(c8mi_em_hybrid* hybrid, c8mi_radio* radio);

/* One shower (HybridCascade::run): begins the shower on both handles
 * (c8mi_em_begin_shower / c8mi_em_scalar_begin_shower), puts the primaries on
 * the host stack (LIFO), and runs until the stack, the staged batch and the
 * device pending queues are empty. History IDs come from one counter starting
 * at first_history_id: children of host steps take the next IDs, device calls
 * reserve ranges (EmRouter.inl:640-679). Ledger: the host accumulators hold
 * the primaries' energy as injected, less what the device calls kept. */
int c8mi_em_hybrid_run(c8mi_em_hybrid* hybrid, const c8mi_particle_state* primaries, uint64_t count,
                       uint64_t shower_id, uint64_t first_history_id, c8mi_em_hybrid_result* out);

/* The CPU prefix alone (the shower up to its first device call): begins the
 * shower on the stepper and returns the batch the router would hand to the
 * device (*staged_count particles; up to `capacity` copied), with
 * *next_history_id the counter after the prefix. For a pool: enqueue the
 * batch in a slot begun with that history ID (c8mi_em_slot_begin). The
 * stepper's ledger holds the primaries' energy as injected less the batch's. */
int c8mi_em_hybrid_prefix(c8mi_em_scalar* scalar, const c8mi_em_hybrid_config* config,
                          const c8mi_particle_state* primaries, uint64_t count, uint64_t shower_id,
                          uint64_t first_history_id, c8mi_particle_state* staged, uint64_t capacity,
                          uint64_t* staged_count, uint64_t* next_history_id, double* tracks,
                          uint64_t track_capacity, uint64_t* track_count, c8mi_em_hybrid_result* out);

/* The CPU prefixes of `count` showers (primaries[k], showerId shower_ids[k],
 * first history ID first_history_id) on `threads` host threads, one stepper
 * per thread built from the given configuration, tables and environment: the
 * batches concatenated in shower order into `staged` (shower k's at
 * offsets[k] .. offsets[k + 1], count + 1 offsets; more than `capacity` ->
 * length_error), each shower's counter after its prefix, its host ledger
 * (NULL skips) and host histograms ([k][4][profile_bins] as
 * c8mi_em_download_profile; NULL skips); *totals sums the counts. */
int c8mi_em_hybrid_prefixes(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                            const c8mi_environment* environment, const c8mi_em_hybrid_config* hybrid,
                            const c8mi_particle_state* primaries, uint64_t count, const uint64_t* shower_ids,
                            uint64_t first_history_id, uint32_t threads, c8mi_particle_state* staged,
                            uint64_t capacity, uint64_t* offsets, uint64_t* next_history_ids,
                            c8mi_em_ledger* ledgers, double* histograms, c8mi_em_hybrid_result* totals);

#ifdef __cplusplus
}
#endif

#endif /* C8MI_HYBRID_H */
