/*
 * c8mi oracle — TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C FP64 restatement of the CORSIKA8-kokkos algorithms on the EM/radio hot
 * path. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker / CPU baseline. The product
 * (corsika8-kokkos_amd/) never links or calls it.
 *
 * Every function cites the reference file:line it follows. Build: oracle/Makefile
 * (gcc -O2 -ffp-contract=off, so no fused multiply-add changes the rounding).
 */
#ifndef C8_ORACLE_H
#define C8_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "c8mi/c8mi_tables.h"
#include "c8mi/c8mi_em.h"
#include "c8mi/c8mi_hybrid.h"
#include "c8mi/c8mi_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- radio (RadioProjection.inl) ---------------- */

/* Observer view after RadioAccumulator::makeObservers
 * (corsika/detail/kokkos/radio/RadioAccumulator.inl:591-617). */
typedef struct or_radio_observer {
  double position[3];
  double start_time;
  double duration;
  double sample_rate;
  double fixed_scale;
  double inverse_fixed_scale;
  uint64_t number_of_bins;
  uint64_t waveform_offset;
} or_radio_observer;

typedef struct or_radio_counters {
  uint64_t coreas_contributions;
  uint64_t zhs_contributions;
  uint64_t zhs_subtracks;
  uint64_t valid_tracks;
  uint64_t fixed_point_overflows;
} or_radio_counters;

/* Builds the integrated table and the edge slopes from refractivity samples,
 * restating make_radio_config (corsika/detail/kokkos/radio/atmosphere/
 * FlatAtmosphereConfig.inl:31-73). integrated must hold n entries. */
void oracle_radio_build_flat_table(const double* refractivity, uint64_t n,
                                   double inverse_step, double* integrated,
                                   double slopes[4]);

/* Projects tracks onto every observer with CoREAS (fields) and/or ZHS
 * (vector potentials); both algorithms share the observer list, as in the
 * reference's fused launch (RadioAccumulator.inl:212-214, 343-354).
 * tracks: [9][stride] compact form (c8mi_types.h).
 * *_fixed: 3*total_bins int64 (x block, y block, z block) or NULL.
 * *_float: 3*total_bins double, unquantised FP64 sums, or NULL.
 * Fixed-point scales follow makeObservers (RadioAccumulator.inl:591-617).
 * threads <= 1 runs serially; otherwise OpenMP over tracks with private
 * waveforms merged at the end (the reference's OpenMP design,
 * RadioAccumulator.inl:718-791). Returns 0 on success. */
int oracle_radio_project(const c8mi_radio_propagation* propagation,
                         uint32_t zhs_refinement, double fixed_point_limit,
                         const c8mi_radio_observer* observers, uint64_t observer_count,
                         int coreas, int zhs, const double* tracks, uint64_t stride,
                         uint64_t track_count, int64_t* coreas_fixed,
                         double* coreas_float, int64_t* zhs_fixed, double* zhs_float,
                         or_radio_counters* counters, int threads);

/* Single-function probes used by the tests to pin the restatement. */
double oracle_radio_nearest_bin(double scaled_time);
double oracle_radio_doppler(double index, const double beta[3], const double emit[3],
                            int device_variant);
/* propagate(): out[0]=time, [1]=n_source, [2]=n_destination, [3]=distance,
 * [4..6]=emit. */
void oracle_radio_propagate(const c8mi_radio_propagation* propagation,
                            const double source[3], const double destination[3],
                            double out[7]);
void oracle_radio_separate_same_bin(double* start_time, double* end_time,
                                    double grid_resolution, int signed_order);

/* Decodes a fixed-point block into V/m (CoREAS) or V s/m (ZHS),
 * RadioAccumulator::downloadSet (RadioAccumulator.inl:850-899). */
void oracle_radio_decode(const int64_t* fixed, uint64_t total_bins,
                         const c8mi_radio_observer* observers, uint64_t observer_count,
                         double fixed_point_limit, int zhs_potential, double* out);

/* ---------------- Philox (Philox.inl:15-57) ---------------- */
void oracle_philox4x32_10(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]);
void oracle_random_words(uint64_t seed, uint64_t shower, uint64_t history, uint64_t step,
                         uint32_t process, uint64_t draw, uint32_t out[4]);
double oracle_uniform_open01(uint64_t seed, uint64_t shower, uint64_t history, uint64_t step,
                             uint32_t process, uint64_t draw, uint32_t lane);

/* ---------------- native tables (ProposalNativeQueries.inl) ---------------- */
void oracle_native_queries(const c8mi_native_table_set* table, const c8mi_native_query* queries,
                           uint64_t count, c8mi_native_result* out);
double oracle_native_physical_loss(const c8mi_native_table_set* table, uint32_t column,
                                   double energy_mev, double transformed_loss);
double oracle_native_transformed_loss(const c8mi_native_table_set* table, uint32_t column,
                                      double energy_mev, double physical_loss);

/* ---------------- EM transport (em_oracle.c) ---------------- */

/* Fixed-point counter layout shared with the device ledger: energy terms
 * (int64, profile_energy_scale units per GeV) first, then plain counts. */
enum {
  OR_EMC_INJECTED = 0,
  OR_EMC_CONTINUOUS = 1,
  OR_EMC_CUT_DEPOSIT = 2,
  OR_EMC_CUT_REST_MASS = 3,
  OR_EMC_OBSERVED = 4,
  OR_EMC_ESCAPED = 5,
  OR_EMC_OVERLAP = 6,
  OR_EMC_MASS_CORRECTION = 7,
  OR_EMC_MEDIUM_REST_MASS = 8,
  OR_EMC_PHOTOELECTRIC_BINDING = 9,
  OR_EMC_FALLBACK_ENERGY = 10,
  OR_EMC_THINNING_ENERGY = 11,
  OR_EMC_PENDING = 12,
  OR_EMC_STEPS = 13,
  OR_EMC_PHOTON_CUTS = 14,
  OR_EMC_LEPTON_LIMIT0 = 15, /* 8 entries */
  OR_EMC_THIN_HILLAS = 23,
  OR_EMC_THIN_STATISTICAL = 24,
  OR_EMC_THIN_DISCARDED = 25,
  OR_EMC_DEPOSITED_STEPS = 26,
  OR_EMC_OVERFLOWS = 27,
  OR_EMC_INVALID = 28,
  OR_EMC_FALLBACK_REASON0 = 29, /* 32 entries */
  OR_EMC_NATIVE_MUON_CUT = 61,  /* signed E - m of photon mu-pair muons cut at their first step */
  OR_EM_COUNTERS = 64
};

typedef c8mi_em_run_result or_em_stats;

/* Runs one shower to completion (or maximum_fronts rounds) with the reference's
 * call schedule as c8mi_em_run runs it: per round the shower's photon front,
 * then its lepton front, when its schedule runs one (a photon call of at most
 * 16 fronts, then a lepton call of at most 1024, each with its own reserved
 * history range starting from first_history_id; fronts of at least 256
 * particles stably sorted by oracle_em_bucket_key; below-minimum-batch
 * checkpoint config->resident_minimum_batch). stats->next_history_id is the
 * next ID the shower's counter would reserve. Histograms hold profile_bins
 * int64 each; counters OR_EM_COUNTERS; observations / fallbacks up to the
 * configured capacities (may be NULL). */
int oracle_em_run_shower(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                         const c8mi_environment* environment, uint64_t shower_id,
                         uint64_t first_history_id, const c8mi_particle_state* primaries,
                         uint64_t primary_count, uint64_t maximum_fronts, int64_t* photons,
                         int64_t* electrons, int64_t* positrons, int64_t* energy_loss,
                         int64_t* counters, c8mi_particle_state* observations,
                         c8mi_em_fallback* fallbacks, or_em_stats* stats);

/* oracle_em_run_shower that also returns the compact radio track of every e+-
 * step, row-major [track_count][9] (the fields of the [9][stride] form of
 * c8mi_types.h), in a buffer the caller releases with oracle_free. */
int oracle_em_run_shower_tracks(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                                const c8mi_environment* environment, uint64_t shower_id,
                                uint64_t first_history_id, const c8mi_particle_state* primaries,
                                uint64_t primary_count, uint64_t maximum_fronts, int64_t* photons,
                                int64_t* electrons, int64_t* positrons, int64_t* energy_loss,
                                int64_t* counters, c8mi_particle_state* observations,
                                c8mi_em_fallback* fallbacks, or_em_stats* stats, double** tracks,
                                uint64_t* track_count);
void oracle_free(void* p);

/* oracle_em_run_shower returning the writers' first-interaction snapshot of
 * the shower (*captured 0: its generation-zero particle reached no final
 * state). */
int oracle_em_first_interaction(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                                const c8mi_environment* environment, uint64_t shower_id,
                                uint64_t first_history_id, const c8mi_particle_state* primaries,
                                uint64_t primary_count, c8mi_em_first_interaction* out, int* captured,
                                or_em_stats* stats);

/* Vertex log of a shower (test probe): one record per lepton vertex attempt
 * (kind 0: the reselection of LeptonVertexSelection.inl:41-199 at the
 * post-loss energy, accepted when process_uniform * rate_start < rate_vertex)
 * and per materialised final state (kind 1), with the kinematic energies and
 * directions of its children before thinning. */
typedef struct or_vertex {
  int32_t kind;
  int32_t pid;          /* parent */
  int32_t process;      /* process id (kind 1) */
  uint32_t accepted;    /* kind 0 */
  uint32_t count;       /* children logged (kind 1) */
  uint32_t pad;
  double energy;        /* parent total energy at the vertex, GeV */
  double mass;          /* projectile mass the kinematics used, GeV */
  double target_mass;   /* target electron mass the kinematics used, GeV (0: none) */
  double direction[3];
  double v;             /* energy fraction */
  double rate_start;    /* kind 0: total rate at the step start */
  double rate_vertex;   /* kind 0: total rate at the vertex */
  double process_uniform;
  int32_t child_pid[3];
  int32_t pad2;
  double child_energy[3];        /* total energies, GeV */
  double child_direction[3][3];
  uint32_t thin_status;          /* two-child thinning of the vertex (kind 1): status, */
  uint32_t keep;                 /* keep mask, */
  double parent_weight;          /* the parent's weight and */
  double child_weight[2];        /* the two children's weights */
} or_vertex;

/* project_profile_grammage (ProfileProjection.inl:12-29) of a point on the
 * configuration's shower axis table. */
double oracle_em_project_grammage(const c8mi_em_config* config, const double position[3]);

/* oracle_em_run_shower with the vertex log, released with oracle_free. */
int oracle_em_run_shower_vertices(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                                  const c8mi_environment* environment, uint64_t shower_id,
                                  uint64_t first_history_id, const c8mi_particle_state* primaries,
                                  uint64_t primary_count, or_vertex** vertices, uint64_t* vertex_count,
                                  or_em_stats* stats);

/* select_discrete_interaction of one particle (InteractionSelection.inl:12-215):
 * status, selected process and component (photons), sampled interaction
 * grammage, total rate and loss fraction. */
typedef struct or_selection {
  int32_t status;
  int32_t process;
  uint64_t component;
  double interaction_grammage;
  double total_rate;
  double energy_fraction;
  double loss_quantile;
  double process_uniform;
  int32_t fallback;     /* 1 when the selection asked for a fallback */
  int32_t fallback_reason;
} or_selection;
int oracle_em_select(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                     const c8mi_environment* environment, uint64_t shower_id,
                     const c8mi_particle_state* particle, or_selection* out);

/* Test hook: the lepton history reservation max(3 n, min(chunk, multiplier *
 * 3 n)) of oracle_em_run_shower (0, 0: the reference's 64 and 2^20). Process
 * wide; not thread safe. */
void oracle_em_set_lepton_reservation(uint64_t multiplier, uint64_t chunk);

/* wavefront_bucket_key (WavefrontBucketing.inl:72-79). */
uint64_t oracle_em_bucket_key(const c8mi_particle_state* p);

/* oracle_em_run_shower over shower_count single-primary showers on `threads`
 * host threads (shower k: primaries[k], shower_id first_shower + k, first
 * history 2); per-shower statistics only. */
int oracle_em_run_batch(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                        const c8mi_environment* environment, uint64_t first_shower,
                        const c8mi_particle_state* primaries, uint64_t shower_count,
                        uint64_t maximum_fronts, int threads, or_em_stats* stats);

/* oracle_em_run_batch keeping each shower's histograms (4 x profile_bins
 * int64 per shower) and counters (OR_EM_COUNTERS per shower). */
int oracle_em_run_batch_profiles(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                                 const c8mi_environment* environment, uint64_t first_shower,
                                 const c8mi_particle_state* primaries, uint64_t shower_count, int threads,
                                 or_em_stats* stats, int64_t* hist, int64_t* counters);

/* Host handoff of one shower (HybridCascade::run, HybridCascade.inl:55-175,
 * with EmRouter::{canRoute, stage, advanceOneWavefrontAndReturn},
 * EmRouter.inl:64-423): LIFO host stack, host steps of the particles the
 * router returns (at most maximum_expansion_rounds consecutive expansion
 * rounds of batches below the minimum), device calls of the staged batch as
 * IEmBackend wavefront calls (device pending prefix, then host particles;
 * remaining particles back to the batch). Host steps are this oracle's step
 * of one source with history IDs from the shower's counter. Each side's
 * histograms (4 x profile_bins, caller's), counters, records (caller's
 * arrays of the configured capacities; counts returned) and compact tracks
 * (oracle-allocated [track_count][9], released with oracle_free).
 * prefix_only: stop at the first device call and return its batch in
 * *staged (oracle_free), the host ledger less the batch; -2 when particles
 * are left on the host stack at that point. limits NULL: unlimited. */
typedef struct or_hybrid_side {
  int64_t* hist;
  int64_t counters[OR_EM_COUNTERS];
  c8mi_particle_state* observations;
  uint64_t observation_count;
  c8mi_em_fallback* fallbacks;
  uint64_t fallback_count;
  double* tracks;
  uint64_t track_count;
  c8mi_em_first_interaction first;
  uint64_t first_captured;
} or_hybrid_side;
int oracle_em_run_hybrid(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                         const c8mi_environment* environment, uint64_t shower_id, uint64_t first_history_id,
                         const c8mi_particle_state* primaries, uint64_t primary_count,
                         const c8mi_em_hybrid_config* hc, const c8mi_em_limits* limits, int prefix_only,
                         or_hybrid_side* host, or_hybrid_side* device, c8mi_particle_state** staged,
                         uint64_t* staged_count, c8mi_em_hybrid_result* out);

/* Molière probes (analytical correction functions). */
double oracle_moliere_f1(const c8mi_moliere_snapshot* m, double x);
double oracle_moliere_f2(const c8mi_moliere_snapshot* m, double x);
double oracle_moliere_integral_f1(const c8mi_moliere_snapshot* m, double x);
double oracle_moliere_integral_f2(const c8mi_moliere_snapshot* m, double x);
double oracle_moliere_solve_b(double rhs);
double oracle_moliere_normal_ppf(double p);
int oracle_moliere_evaluate(const c8mi_moliere_snapshot* m, double grammage, double e_mev, double theta,
                            double* density, double* cumulative);
int oracle_moliere_angle(const c8mi_moliere_snapshot* m, double grammage, double e_mev, double u1, double u2,
                         uint32_t* iterations, double* angle);
/* The same with an interpolation table (NULL: analytical functions). */
int oracle_moliere_angle_interpolated(const c8mi_moliere_snapshot* m, const c8mi_moliere_interpolation* t,
                                      double grammage, double e_mev, double u1, double u2, uint32_t* iterations,
                                      double* angle);
void oracle_moliere_interpolated_functions(const c8mi_moliere_snapshot* m, const c8mi_moliere_interpolation* t,
                                           double x, double out[4]);
int oracle_moliere_direction(const double start[3], const double tracked[3], double angle, double au,
                             double out[3]);

int oracle_em_environment_query(const c8mi_environment* e, const double position[3],
                                const double direction[3], int32_t node, int32_t* out_node,
                                double* out_density);
int oracle_em_boundary(const c8mi_environment* e, const double position[3],
                       const double direction[3], int32_t node, double* distance,
                       int32_t* next_node);
int oracle_em_grammage(const c8mi_environment* e, int32_t node, const double position[3],
                       const double direction[3], double distance, double* out);
int oracle_em_arclength(const c8mi_environment* e, int32_t node, const double position[3],
                        const double direction[3], double grammage, double* out);
int oracle_em_pair_lpm(const c8mi_photon_pair_lpm* s, uint64_t comp, double e_mev, double x,
                       double rho, double* out);
int oracle_em_brems_lpm(const c8mi_brems_lpm* s, uint64_t comp, double e_mev, double v,
                        double rho, double* out);
int oracle_em_thinning(const c8mi_em_thinning* c, double pe, double pw, double e1, double e2,
                       double u1, double u2, uint32_t* keep, double* w1, double* w2);
int oracle_em_pair_split(double e_gev, double x, double* normalized, double* decoded);
double oracle_em_sauter_cosine(double e, double u);
int oracle_em_compton_cosines(double e, double v, double* pc, double* ec);
int oracle_em_annihilation_rho(double e, double m, double u, double* rho);

#ifdef __cplusplus
}
#endif

#endif /* C8_ORACLE_H */
