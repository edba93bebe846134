/*
 * c8mi EM C ABI — MI355X replacement for the device side of CORSIKA8-kokkos's
 * electromagnetic backend (corsika/kokkos/em). Every entry point returns a
 * c8mi_status (0 = success, codes as in c8mi_radio.h); on failure
 * c8mi_last_error() returns the message for the calling thread.
 *
 * Device pointers are plain hipMalloc'd addresses (c8mi_device_malloc). All
 * launches are asynchronous on the given HIP stream (NULL = the null stream).
 */
#ifndef C8MI_EM_H
#define C8MI_EM_H

#include "c8mi_radio.h"
#include "c8mi_tables.h"
#include "c8mi_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* RandomNumberKey (corsika/kokkos/em/random/Philox.hpp:27-35), 48 bytes. */
typedef struct c8mi_random_key {
  uint64_t seed;
  uint64_t shower_id;
  uint64_t history_id;
  uint64_t step_id;
  uint32_t process_id;
  uint32_t pad_;
  uint64_t draw_id;
} c8mi_random_key;

/* Device copy of one format-v7 native table set: the storage the reference
 * allocates in BackendInstance::initialize from a ProposalNativeTableSet
 * (corsika/kokkos/em/tables/ProposalNativeTable.hpp:21-56,
 * BackendInstanceImplementation.inl). The host set is deep-copied; bounds of
 * every spline descriptor and index are checked first (invalid_argument). */
typedef struct c8mi_em_tables c8mi_em_tables;

int c8mi_em_tables_create(const c8mi_native_table_set* host_tables, c8mi_em_tables** out);
int c8mi_em_tables_destroy(c8mi_em_tables* tables);
/* The device view: same struct, device pointers (ProposalNativeDeviceView). */
int c8mi_em_tables_view(const c8mi_em_tables* tables, c8mi_native_table_set* device_view);
/* Device bytes held by the tables (proposal_native_table_bytes,
 * ProposalNativeTable.hpp:66-67). */
int c8mi_em_tables_bytes(const c8mi_em_tables* tables, uint64_t* bytes);

/* Batched PROPOSAL-native queries: one device thread per query, results as the
 * reference's query_proposal_native_* functions
 * (corsika/detail/kokkos/em/tables/ProposalNativeQueries.inl:744-1428) return
 * them (status, Newton/bisection iteration counts, value). Both arrays are
 * device memory of `count` entries. */
int c8mi_em_native_queries(const c8mi_em_tables* tables, const c8mi_native_query* device_queries,
                           uint64_t count, c8mi_native_result* device_results, void* stream);

/* Batched Philox4x32-10 draws (random_words / uniform_open01,
 * corsika/detail/kokkos/em/random/Philox.inl:23-57) for `count` device keys:
 * four words and the four open-interval uniforms of each key. Either output
 * may be NULL. */
int c8mi_em_random(const c8mi_random_key* device_keys, uint64_t count, uint32_t* device_words,
                   double* device_uniforms, void* stream);

/* ---- Environment (corsika/kokkos/environment) ------------------------------ */

/* EnvironmentModel::Density (corsika/kokkos/environment/EnvironmentModel.hpp:29-65). */
enum {
  C8MI_DENSITY_HOMOGENEOUS = 1,
  C8MI_DENSITY_SLIDING_PLANAR_EXPONENTIAL = 2, /* height = |x - anchor| */
  C8MI_DENSITY_FLAT_EXPONENTIAL = 3            /* height = (x - anchor) . axis */
};
/* EnvironmentShape::Kind (corsika/kokkos/environment/EnvironmentShape.hpp:27-36). */
enum { C8MI_SHAPE_UNIVERSE = 1, C8MI_SHAPE_SPHERE = 2 };

/* One medium model: rho = rho0 * exp(inverse_scale_length * (height - reference_height)). */
typedef struct c8mi_env_model {
  uint32_t density_kind;
  int32_t medium_id;
  double reference_density_g_per_cm3;
  double scale_length_m; /* signed; 0 for homogeneous */
  double inverse_scale_length_per_m;
  double reference_height_m;
  double anchor_m[3];
  double axis[3];
  double magnetic_field_t[3]; /* uniform B (tesla) of the model's volume; finite */
} c8mi_env_model;

/* One volume of the environment tree (EnvironmentTreeView): its shape, parent,
 * ordered children (a range of c8mi_environment.children) and medium model
 * (-1: no medium, the particle leaves the environment there). */
typedef struct c8mi_env_node {
  uint32_t shape_kind;
  int32_t parent; /* -1 for the root */
  int32_t model;
  uint32_t child_begin;
  uint32_t child_count;
  uint32_t pad_;
  double center_m[3];
  double radius_m;
} c8mi_env_node;

/* Environment snapshot + transport geometry (TransportGeometryConfig): tree,
 * models, observation plane (point, unit normal) and Earth centre. */
typedef struct c8mi_environment {
  const c8mi_env_node* nodes;
  uint32_t node_count;
  uint32_t child_index_count;
  const int32_t* children;
  const c8mi_env_model* models;
  uint32_t model_count;
  uint32_t observation_enabled;
  double observation_point_m[3];
  double observation_normal[3];
  double earth_center_m[3];
  double maximum_magnetic_deflection_rad;
} c8mi_environment;

/* ---- Physics auxiliaries ------------------------------------------------------ */

/* PhotonPairLpmSnapshot (corsika/kokkos/em/models/PhotonPairLpm.hpp:20-40). */
typedef struct c8mi_photon_pair_lpm_component {
  uint64_t component_hash;
  double nuclear_charge;
  double radiation_log_constant;
} c8mi_photon_pair_lpm_component;

typedef struct c8mi_photon_pair_lpm {
  double baseline_mass_density_g_per_cm3;
  double molecular_density_per_cm3;
  double sum_charge;
  double e_lpm_mev;
  double classical_electron_radius_cm;
  double fine_structure_constant;
  uint32_t component_count;
  uint32_t reserved;
  c8mi_photon_pair_lpm_component components[16];
} c8mi_photon_pair_lpm;

/* BremsLpmSnapshot (corsika/kokkos/em/models/BremsLpm.hpp:21-51). */
typedef struct c8mi_brems_lpm_component {
  uint64_t component_hash;
  double nuclear_charge;
  double atomic_mass_number;
  double radiation_log_constant;
} c8mi_brems_lpm_component;

typedef struct c8mi_brems_lpm {
  double baseline_mass_density_g_per_cm3;
  double molecular_density_per_cm3;
  double sum_charge;
  double e_lpm_mev;
  double lepton_mass_mev;
  double electron_mass_mev;
  double muon_mass_mev;
  double classical_electron_radius_cm;
  double fine_structure_constant;
  uint32_t component_count;
  uint32_t reserved;
  c8mi_brems_lpm_component components[16];
  double muon_baseline_mass_density_g_per_cm3;
  double muon_molecular_density_per_cm3;
  double muon_sum_charge;
  double muon_e_lpm_mev;
} c8mi_brems_lpm;

/* MoliereSnapshot (corsika/kokkos/em/models/MoliereSnapshot.hpp:22-55): the
 * per-medium state of Molière multiple scattering, byte-compatible with the
 * reference's record. Components: charge, weight * Z^2, chi_0^2 [MeV^2] and the
 * Coulomb correction 3.76 (alpha Z)^2; c1/c2 the power series of the projected
 * correction functions f1, f2 in x = theta^2 / (chi_c^2 B) for small x,
 * c2_large/s2_large and c1_integral_large their large-x expansions. */
#define C8MI_MOLIERE_MAX_COMPONENTS 16
typedef struct c8mi_moliere_component {
  double nuclear_charge;
  double weight_zz;
  double chi0_squared_mev2;
  double coulomb_correction;
} c8mi_moliere_component;

typedef struct c8mi_moliere_snapshot {
  double particle_mass_mev;
  double electron_mass_mev;
  double fine_structure_constant;
  double avogadro_per_mol;
  double hbar_mev_s;
  double speed_of_light_cm_per_s;
  double z_squared_over_a_average;
  double inverse_weight_zz_sum;
  double euler_mascheroni;
  uint32_t component_count;
  uint32_t maximum_weight_index;
  c8mi_moliere_component components[C8MI_MOLIERE_MAX_COMPONENTS];
  double c1[70];
  double c2[70];
  double c2_large[13];
  double s2_large[13];
  double c1_integral_large[15];
} c8mi_moliere_snapshot;

/* MoliereInterpolationTable (corsika/kokkos/em/tables/MoliereInterpolation.hpp:19-56),
 * byte-compatible with the reference's record and with the table part of its
 * cache file: f1, f2 and the integrals of f1 and f2 (in that order) as cubic
 * polynomials in q = (x - x_k) / h, ascending powers, on the 99 intervals of
 * x in [0, 20]; then the Newton initial-guess corrections, tabulated in
 * reference B (64 nodes on [4.5, 64]), beta^2 (8 on [0.25, 1]) and the
 * Gaussian quantile coordinate (128 on [0, 5]), index (b * 8 + beta) * 128 + c. */
#define C8MI_MOLIERE_INTERVALS 99
#define C8MI_MOLIERE_GUESS_B_NODES 64
#define C8MI_MOLIERE_GUESS_BETA2_NODES 8
#define C8MI_MOLIERE_GUESS_COORDINATE_NODES 128
typedef struct c8mi_moliere_cubic {
  double c[4];
} c8mi_moliere_cubic;

typedef struct c8mi_moliere_interpolation {
  c8mi_moliere_cubic polynomials[4 * C8MI_MOLIERE_INTERVALS];
  double initial_guess_delta[C8MI_MOLIERE_GUESS_B_NODES * C8MI_MOLIERE_GUESS_BETA2_NODES *
                             C8MI_MOLIERE_GUESS_COORDINATE_NODES];
} c8mi_moliere_interpolation;

/* make_moliere_interpolation_table (src/kokkos/em/tables/MoliereInterpolation.cpp:438-460)
 * on the host: interpolating cubic splines with natural ends through f1, f2 and
 * their integrals at the 100 nodes, written per interval in the Hermite form of
 * the reference (:272-310), then the initial-guess corrections by bisection of
 * the mixture CDF (:362-436). The reference builds its splines with the
 * CubicInterpolation library (not part of this build); its cache file holds the
 * exact production table and can be read with c8mi_em_moliere_interpolation_read. */
int c8mi_em_moliere_interpolation_build(const c8mi_moliere_snapshot* snapshot,
                                        c8mi_moliere_interpolation* out);

/* read_interpolation_cache / write_interpolation_cache
 * (MoliereInterpolation.cpp:204-270): the reference's version-1 cache file
 * (header with the snapshot and an FNV-1a checksum of the table, then the
 * table). Reading accepts the reference's magic "C8MOLI01" and this build's
 * "C8MIMO01", and fails (status 1, invalid argument) unless the file is
 * complete, its checksum and snapshot match and every entry is finite.
 * Writing stamps "C8MIMO01" (the table may be this build's natural-spline
 * restatement, which the reference's reader must not take for its own), goes
 * through a temporary file in the same directory that is synced, read back and
 * compared before it is renamed over `path`; on failure (status 4) `path` is
 * untouched. */
int c8mi_em_moliere_interpolation_read(const char* path, const c8mi_moliere_snapshot* snapshot,
                                       c8mi_moliere_interpolation* out);
int c8mi_em_moliere_interpolation_write(const char* path, const c8mi_moliere_snapshot* snapshot,
                                        const c8mi_moliere_interpolation* table);

/* EmThinningConfig (corsika/kokkos/em/algorithms/EmThinning.hpp:25-30). */
typedef struct c8mi_em_thinning {
  double threshold_gev;
  double maximum_weight;
  uint32_t enabled;
  uint32_t erase_zero_weight;
} c8mi_em_thinning;

/* ---- EM backend (IEmBackend, corsika/kokkos/em/execution/IEmBackend.hpp:25-88) -- */

enum { C8MI_PROFILE_CROSSING_BOTH = 0, C8MI_PROFILE_CROSSING_FORWARD = 1 };

/* BackendConfig + ShowerConfig subset of the device path. The projection table
 * (ProfileProjectionView) maps a position to grammage along the shower axis:
 * axis_grammage[k] is the grammage at axis_start + k * axis_step * axis_direction. */
typedef struct c8mi_em_config {
  uint64_t random_seed;
  double em_transport_cut_mev;   /* e+- kinetic and photon energy cut */
  double muon_transport_cut_mev; /* muon kinetic cut (muons fall back to the host) */
  uint64_t queue_capacity;       /* particles per species and buffer */
  uint64_t observation_capacity; /* stored observation-surface particles */
  uint64_t fallback_capacity;    /* stored host fallback requests */
  c8mi_photon_pair_lpm photon_pair_lpm;
  c8mi_brems_lpm brems_lpm;
  c8mi_em_thinning thinning;
  /* Molière multiple scattering of e+- steps (MoliereStep.inl:74-101): enabled
   * when moliere_enabled is 1; the snapshot is the electron one of the medium. */
  c8mi_moliere_snapshot electron_moliere;
  uint32_t moliere_enabled;
  /* Newton iterations per projected angle in the step kernel (0: 12); a sample
   * still unconverged continues on a workgroup of its own. Results do not
   * depend on it (launch tuning only). */
  uint32_t moliere_fast_iterations;
  /* The interpolation table of that snapshot (the reference backend always
   * builds one, BackendInstanceImplementation.inl:469): f1, f2 and their
   * integrals from the table for x <= 20 and the tabulated Newton initial
   * guess. NULL: the analytical functions everywhere. Host memory, copied at
   * create. */
  const c8mi_moliere_interpolation* electron_moliere_interpolation;
  uint32_t profile_bins;
  uint32_t profile_crossing_mode;
  double profile_bin_width_g_per_cm2;
  double profile_energy_loss_threshold_g_per_cm2;
  double profile_weight_scale;     /* fixed-point units per unit weight */
  double profile_energy_scale;     /* fixed-point units per GeV */
  double axis_start_m[3];
  double axis_direction[3];
  double axis_step_m;
  const double* axis_grammage_g_per_cm2;
  uint64_t axis_support_count;
  /* Concurrent showers (pool slots, 1..4096; 0 means 1). Each slot has its
   * own Philox shower key, history counter, projection axis, profile, ledger,
   * observation_capacity observations and fallback_capacity fallbacks; the
   * particle queues (queue_capacity per species and buffer) are shared. */
  uint32_t shower_slots;
  uint32_t reserved_;
  /* IEmBackend::minimumBatchSize (BackendConfig::minBatchSize_,
   * corsika/kokkos/em/execution/BackendConfig.hpp:26; 0 selects its default
   * 16384): the handoff batch a host router stages towards
   * (EmRouter.inl:246-260). */
  uint64_t minimum_batch_size;
  /* Below-minimum-batch checkpoint of the device-resident drivers (c8mi_em_run,
   * c8mi_em_run_rounds, c8mi_em_serve): a shower's call of one species ends
   * after its first front once that shower's next front would hold fewer
   * particles (PhotonCascade.inl:296-300); the rest waits for the shower's
   * next call of the species. 0 or 1: never. */
  uint64_t resident_minimum_batch;
} c8mi_em_config;

/* Shower axis of one pool slot (the ProfileProjectionView of its ShowerConfig):
 * grammage[k] at start_m + k * step_m * direction. */
typedef struct c8mi_em_shower_axis {
  double start_m[3];
  double direction[3];
  double step_m;
  const double* grammage_g_per_cm2;
  uint64_t support_count;
} c8mi_em_shower_axis;

/* Progress of one pool slot. */
typedef struct c8mi_em_slot_state {
  uint64_t shower_id;
  uint64_t pending_photons; /* queued on the device or staged on the host */
  uint64_t pending_leptons;
  uint64_t next_history_id;
  uint64_t photon_fronts;   /* fronts in which the shower had photons */
  uint64_t lepton_fronts;
  uint64_t photon_steps;
  uint64_t lepton_steps;
  uint64_t photon_interactions;
  uint64_t lepton_interactions;
  uint64_t secondaries;
  uint64_t observations;
  uint64_t fallbacks;
  uint32_t active;          /* begun and not ended */
  uint32_t pad_;
} c8mi_em_slot_state;

/* Per-call statistics (PhotonCascadeResult / LeptonCascadeResult counters). */
typedef struct c8mi_em_run_result {
  uint64_t photon_fronts;
  uint64_t lepton_fronts;
  uint64_t photon_steps;    /* transport records, as PhotonCascade.inl:499 counts them */
  uint64_t lepton_steps;
  uint64_t photon_interactions;
  uint64_t lepton_interactions;
  uint64_t secondaries;
  uint64_t observations;
  uint64_t fallbacks;
  uint64_t pending_photons;
  uint64_t pending_leptons;
  uint64_t next_history_id;
  uint64_t max_front;       /* largest single front processed */
  /* The reference's call schedule (EmRouter.inl:368-407): per shower, a photon
   * call of at most 16 fronts, then a lepton call of at most 1024 fronts, each
   * with its own reserved history range. */
  uint64_t photon_calls;    /* calls that reserved a history range */
  uint64_t lepton_calls;
  uint64_t bucketed_fronts;     /* (shower, front) pairs of >= 256 particles, stably
                                   sorted by the bucket key (WavefrontBucketing.inl:72-79) */
  uint64_t bucketed_particles;
  uint64_t below_minimum_batch_checkpoints;
  uint64_t history_range_checkpoints;
  uint64_t reserved_history_ids;
} c8mi_em_run_result;

/* Energy ledger of one shower (ProfileCounters / ProfileResult ledger terms,
 * corsika/kokkos/output/ProfileCounters.hpp; closure ShowerReport.inl:1236-1276).
 * Weighted energies in GeV, decoded from the device's int64 fixed point. */
typedef struct c8mi_em_ledger {
  double injected_energy;               /* weighted total energy enqueued, less the
                                           remaining particles wavefront calls returned */
  double continuous_deposit;            /* dE/dX along lepton steps */
  double cut_deposit;                   /* kinetic (photon: total) energy at particle cuts */
  double cut_rest_mass;                 /* lepton rest mass at particle cuts */
  double observed_energy;               /* total energy reaching the observation surface */
  double escaped_energy;                /* total energy leaving the environment */
  double observation_cut_overlap;       /* energy counted both as observed and cut */
  double mass_convention_correction;    /* transport vs native mass of kept leptons */
  double medium_rest_mass_input;        /* atomic electrons set in motion */
  double photoelectric_binding;         /* photoelectric binding energy */
  double fallback_energy;               /* handed to the host in fallback requests */
  double thinning_discarded_energy;     /* weight-energy removed by thinning */
  double pending_energy;                /* still queued on the device */
  uint64_t steps;
  uint64_t photon_cuts;
  uint64_t lepton_limits[8];
  uint64_t thinning_hillas_vertices;
  uint64_t thinning_statistical_vertices;
  uint64_t thinning_particles_discarded;
  uint64_t deposited_steps;
  uint64_t fixed_point_overflows;
  uint64_t invalid_records;
  uint64_t fallback_reasons[32];
  double native_muon_pair_cut_energy;   /* signed E - m of photon μ-pair muons made at or
                                           below their transport mass and cut at their
                                           first step (weightedNativeMuonPairCutEnergy_,
                                           a sink of the muon energy loss) */
} c8mi_em_ledger;

/* FallbackRequest subset: the particle at the failing stage and the reason
 * (FallbackReason, corsika/kokkos/em/FallbackRequest.hpp:22-65). */
/* PhotonCascadeResult / LeptonCascadeResult of one IEmBackend wavefront call
 * (corsika/kokkos/em/execution/{Photon,Lepton}CascadeResult.hpp). */
typedef struct c8mi_em_wavefront_result {
  uint64_t input_particles;      /* device pending prefix + host particles */
  uint64_t wavefronts;
  uint64_t transport_records;    /* steps (PhotonCascade.inl:499) */
  uint64_t interactions;
  uint64_t secondaries;          /* children given history IDs */
  uint64_t next_secondary_history_id;
  uint64_t peak_resident;        /* largest front of the call */
  uint64_t remaining;            /* particles of the species left when the call
                                    ended (remainingPhotons_ / remainingLeptons_),
                                    c8mi_em_download_remaining */
  uint64_t other_species_pending; /* the other species' device pending queue */
  uint64_t bucketed_fronts;      /* wavefrontBucketingBatches_ (fronts of >= 256) */
  uint64_t bucketed_particles;
  uint64_t small_fronts;         /* wavefrontBucketingSmallBatches_ */
  uint64_t small_particles;
  uint64_t observations;         /* recorded during the call */
  uint64_t fallbacks;
  uint32_t completed;            /* no particle of the species left */
  uint32_t below_minimum_batch_checkpoint;
  uint32_t history_range_exhausted;
  uint32_t pad_;
} c8mi_em_wavefront_result;

/* IEmBackend::minimumBatchSize and maximumResident{Photon,Lepton,Input}BatchSize
 * (IEmBackend.hpp:31-35). */
typedef struct c8mi_em_limits {
  uint64_t minimum_batch_size;
  uint64_t maximum_resident_photon_batch;
  uint64_t maximum_resident_lepton_batch;
  uint64_t maximum_resident_input_batch;
  uint64_t queue_capacity;
  uint32_t shower_slots;
  uint32_t pad_;
} c8mi_em_limits;

/* The transport counters of IEmBackend::statistics (BackendStatistics,
 * corsika/kokkos/em/execution/BackendStatistics.hpp:20-150), accumulated since
 * c8mi_em_create over every call and slot. */
typedef struct c8mi_em_statistics {
  uint64_t shower_ordinal;        /* showers begun (begin_shower, slot starts, served requests) */
  uint64_t wavefronts;
  uint64_t photon_wavefronts;
  uint64_t lepton_wavefronts;
  uint64_t particles_enqueued;
  uint64_t particles_advanced;    /* transport records */
  uint64_t particles_produced;    /* children given history IDs */
  uint64_t interactions_selected;
  uint64_t bucketed_fronts;
  uint64_t bucketed_particles;
  uint64_t small_fronts;
  uint64_t small_particles;
  uint64_t ordered_fronts;        /* fronts that needed the ordering step */
  uint64_t photon_calls;
  uint64_t lepton_calls;
  uint64_t below_minimum_batch_checkpoints;
  uint64_t history_range_checkpoints;
  uint64_t reserved_history_ids;
  uint64_t remaining_downloaded; /* particles returned by wavefront calls */
  uint64_t host_to_device_bytes;
  uint64_t device_to_host_bytes;
  uint64_t parked_fronts;         /* ordered fronts with showers sitting them out (partitioned behind) */
  uint64_t parked_particles;      /* entries of those showers carried over unchanged */
} c8mi_em_statistics;

typedef struct c8mi_em_fallback {
  c8mi_particle_state particle;
  int32_t reason;
  int32_t process_id;
  uint32_t stage; /* 1 selection, 2 transport, 3 vertex, 4 final state */
  uint32_t pad_;
} c8mi_em_fallback;

/* FirstInteractionRecord (corsika/kokkos/em/FirstInteractionRecord.hpp:22-27):
 * the first final state of the shower's generation-zero particle, its
 * children before thinning, each with the parent's weight (the writers'
 * snapshot, FirstInteractionRecord.hpp:15-19; PhotonFront.inl:95-111). Children
 * carry the parent's history ID (the IDs assigned afterwards are not part of
 * the snapshot). secondary_count entries are valid (1 or 2). */
typedef struct c8mi_em_first_interaction {
  c8mi_particle_state parent_at_vertex;
  c8mi_particle_state secondaries[3];
  int32_t process_id;
  uint32_t secondary_count;
} c8mi_em_first_interaction;

typedef struct c8mi_em c8mi_em;

/* EmBackend::initialize (corsika/kokkos/em/execution/EmBackend.hpp:29-31):
 * copies tables, environment and configuration to the device. */
int c8mi_em_create(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                   const c8mi_environment* environment, void* stream, c8mi_em** out);
int c8mi_em_destroy(c8mi_em* em);

/* IEmBackend::beginShower: clears queues, profile, ledger, observations and
 * fallbacks; sets the Philox shower identity and the first secondary history ID. */
int c8mi_em_begin_shower(c8mi_em* em, uint64_t shower_id, uint64_t first_history_id);

/* IEmBackend::canTransport (photon or e+-, finite, inside the energy domain). */
int c8mi_em_can_transport(const c8mi_em* em, const c8mi_particle_state* particle, int* out);

/* Adds host particles to the device queues (photons and e+- split by PID). */
int c8mi_em_enqueue(c8mi_em* em, const c8mi_particle_state* particles, uint64_t count);

/* IEmBackend::runPhotonWavefront (IEmBackend.hpp:51-53; run_photon_cascade,
 * corsika/detail/kokkos/em/execution/PhotonCascade.inl:177-574) for the shower
 * begun with c8mi_em_begin_shower. The call's input is the device pending
 * photon prefix (photons the lepton fronts appended) followed by `particles`
 * (photons only). At most `maximum_wavefronts` fronts run; each front of at
 * least 256 photons is first stably sorted by the bucket key
 * (WavefrontBucketing.inl:72-79); children take history IDs from
 * first_secondary_history_id on, in front order; after the first front the
 * call stops when fewer than `minimum_resident_batch` photons remain. Charged
 * children stay in the device pending lepton queue. Photons left when the call
 * ends are downloaded (c8mi_em_download_remaining) and no longer queued.
 * Errors as the reference: zero first ID, wavefronts or batch -> 1
 * (invalid_argument); more input than the queue holds -> 3 (length_error);
 * other pool slots active -> 2 (logic_error). */
int c8mi_em_run_photon_wavefront(c8mi_em* em, const c8mi_particle_state* particles,
                                 uint64_t count, uint64_t first_secondary_history_id,
                                 uint64_t maximum_wavefronts, uint64_t minimum_resident_batch,
                                 c8mi_em_wavefront_result* out);
/* IEmBackend::runLeptonWavefront (IEmBackend.hpp:55-58; run_lepton_cascade,
 * LeptonCascade.inl:86-565): as above for e+-, with the history range
 * [first_secondary_history_id, secondary_history_id_limit_exclusive): before
 * each front the call stops when the front holds more than a third of the IDs
 * left (historyRangeExhausted_, LeptonCascade.inl:248-252). Photons join the
 * device pending photon queue. limit <= first -> 1 (invalid_argument). */
int c8mi_em_run_lepton_wavefront(c8mi_em* em, const c8mi_particle_state* particles,
                                 uint64_t count, uint64_t first_secondary_history_id,
                                 uint64_t maximum_wavefronts,
                                 uint64_t secondary_history_id_limit_exclusive,
                                 uint64_t minimum_resident_batch, c8mi_em_wavefront_result* out);
/* The remaining particles of the last wavefront call (remainingPhotons_ /
 * remainingLeptons_): up to `capacity` copied into `out`; *count the total. */
int c8mi_em_download_remaining(c8mi_em* em, c8mi_particle_state* out, uint64_t capacity,
                               uint64_t* count);

/* IEmBackend limits (IEmBackend.hpp:31-35). */
int c8mi_em_get_limits(const c8mi_em* em, c8mi_em_limits* out);
/* IEmBackend::statistics (IEmBackend.hpp:73): cumulative transport counters. */
int c8mi_em_get_statistics(const c8mi_em* em, c8mi_em_statistics* out);

/* The device-resident driver: rounds of the reference's call schedule over
 * every active slot (EmRouter.inl:339-416 with the device pending queues, as
 * the router runs a shower whose fronts all stay on the device): per shower a
 * photon call of at most 16 fronts, then a lepton call of at most 1024 fronts,
 * each reserving its history range from the shower's counter
 * (first_history_id of the begin call; photons 2 n per front of the call,
 * leptons max(3 n, min(2^20, 192 n)), EmRouter.inl:640-679), with the
 * bucketing, the call-input order (pending, then remaining, then host) and the
 * exhaustion and resident_minimum_batch checkpoints of the wavefront calls.
 * Runs until nothing is pending or `maximum_rounds` rounds (a photon front,
 * then a lepton front) have run. out->next_history_id: the next ID slot 0's
 * counter would reserve. */
int c8mi_em_run(c8mi_em* em, uint64_t maximum_rounds, c8mi_em_run_result* out);

/* IEmBackend::pendingPhotonCount / pendingLeptonCount. */
int c8mi_em_pending_counts(const c8mi_em* em, uint64_t* photons, uint64_t* leptons);

/* IEmBackend::downloadProfile: five longitudinal histograms of profile_bins
 * entries each (photons, electrons, positrons, e+- energy loss [GeV], and
 * weighted particle counts are decoded) and the ledger. Any output may be NULL. */
int c8mi_em_download_profile(c8mi_em* em, double* photons, double* electrons,
                             double* positrons, double* energy_loss, c8mi_em_ledger* ledger);

/* Observation-surface particles (ObservationRecord) and host fallback requests
 * of the current shower; *count receives the total (which may exceed capacity). */
int c8mi_em_download_observations(c8mi_em* em, c8mi_particle_state* out, uint64_t capacity,
                                  uint64_t* count);
int c8mi_em_download_fallbacks(c8mi_em* em, c8mi_em_fallback* out, uint64_t capacity,
                               uint64_t* count);
/* IEmBackend::downloadFirstInteractionSnapshot (IEmBackend.hpp:87): *captured
 * 0 when the shower's generation-zero particle has had no final state on the
 * device (out then untouched). */
int c8mi_em_download_first_interaction(c8mi_em* em, c8mi_em_first_interaction* out, int* captured);

/* ---- Shower pool: several showers in the same device fronts ----
 * No counterpart in IEmBackend, which holds one shower per backend; this is the
 * device side of a serving loop that keeps many showers resident. Particles of
 * all active slots share the queues, sorted by slot; compaction is stable
 * within each slot and history IDs are drawn per slot, so every shower's
 * Philox draws, records, profile and ledger are those of the same shower run
 * alone. The single-shower entry points above act on slot 0. */

/* Starts a shower in an idle slot (no pending particles): clears its profile,
 * ledger and records, sets its Philox shower key and first secondary history
 * ID and its projection axis (NULL: the axis of the configuration). */
int c8mi_em_slot_begin(c8mi_em* em, uint32_t slot, uint64_t shower_id,
                       uint64_t first_history_id, const c8mi_em_shower_axis* axis);

/* Stages particles of one slot; they join the device queues, sorted by slot,
 * at the start of the next front (of either species). */
int c8mi_em_slot_enqueue(c8mi_em* em, uint32_t slot, const c8mi_particle_state* particles,
                         uint64_t count);

/* Runs up to `maximum_rounds` rounds (a photon front, then a lepton front, each
 * over all slots) or until every queue is empty; *out sums over the slots. */
int c8mi_em_run_rounds(c8mi_em* em, uint64_t maximum_rounds, c8mi_em_run_result* out);

int c8mi_em_slot_state_get(const c8mi_em* em, uint32_t slot, c8mi_em_slot_state* out);

/* Per-slot forms of the downloads above. */
int c8mi_em_slot_download_profile(c8mi_em* em, uint32_t slot, double* photons, double* electrons,
                                  double* positrons, double* energy_loss, c8mi_em_ledger* ledger);
int c8mi_em_slot_download_observations(c8mi_em* em, uint32_t slot, c8mi_particle_state* out,
                                       uint64_t capacity, uint64_t* count);
int c8mi_em_slot_download_fallbacks(c8mi_em* em, uint32_t slot, c8mi_em_fallback* out,
                                    uint64_t capacity, uint64_t* count);
int c8mi_em_slot_download_first_interaction(c8mi_em* em, uint32_t slot, c8mi_em_first_interaction* out,
                                            int* captured);

/* Marks a slot idle. A slot that still has pending particles is refused
 * (logic error): run it to completion first. */
int c8mi_em_slot_end(c8mi_em* em, uint32_t slot);

/* Host-side pending count of a slot (no device call): particles queued on the
 * device plus particles staged on the host. */
int c8mi_em_slot_pending(const c8mi_em* em, uint32_t slot, uint64_t* photons, uint64_t* leptons);

/* ---- Serving loop over the pool ----
 * A stream of shower requests is admitted into free slots as they arrive and
 * as the admitted showers' footprints (estimated peak queue occupancy) fit the
 * budget; every round steps all admitted showers together (continuous
 * batching). A pool of one slot is the one-shower-at-a-time form. */
typedef struct c8mi_em_request {
  c8mi_particle_state primary;
  uint64_t shower_id;
  uint64_t first_history_id;
  double arrival_s;            /* offset from the start of the serve call, non-decreasing */
  double footprint_particles;  /* admission estimate of the shower's peak queue occupancy */
  c8mi_em_shower_axis axis;    /* grammage_g_per_cm2 NULL: the configuration's axis */
  /* Per-shower radio (c8mi_em_enable_slot_radio): the shower's observers
   * (none: no radio) and where its raw fixed-point sums go when it ends,
   * [3][sum of the observers' bins] each (either may be NULL). */
  const c8mi_radio_observer* radio_observers;
  uint64_t radio_observer_count;
  int64_t* radio_coreas;
  int64_t* radio_zhs;
  /* A batch the host handed off (the CPU prefix of the shower,
   * c8mi_em_hybrid_prefix): with particle_count > 0 these particles are
   * enqueued at admission instead of `primary`, and first_history_id is the
   * prefix's counter. Read during the serve call only. */
  const c8mi_particle_state* particles;
  uint64_t particle_count;
} c8mi_em_request;

typedef struct c8mi_em_serve_options {
  double footprint_budget_particles; /* admitted footprints stay below this (one shower always fits) */
  uint32_t realtime;                 /* 1: admit at the wall-clock arrival; 0: ignore arrival times */
  uint32_t pad_;
} c8mi_em_serve_options;

typedef struct c8mi_em_request_result {
  double admitted_s;  /* offsets from the start of the serve call */
  double finished_s;
  uint64_t photon_steps;
  uint64_t lepton_steps;
  uint64_t photon_interactions;
  uint64_t lepton_interactions;
  uint64_t secondaries;
  uint64_t next_history_id;
  uint64_t observations;
  uint64_t fallbacks;
  uint32_t slot;
  uint32_t pad_;
  c8mi_em_ledger ledger;
  c8mi_radio_counters radio_counters; /* per-shower radio only */
} c8mi_em_request_result;

/* Runs the whole request stream to completion; *totals sums the fronts. The
 * pool must hold no pending particles on entry (finished showers are ended);
 * it is idle on return. */
int c8mi_em_serve(c8mi_em* em, const c8mi_em_request* requests, uint64_t count,
                  const c8mi_em_serve_options* options, c8mi_em_request_result* results,
                  c8mi_em_run_result* totals);

/* ---- Radio from the lepton fronts ----
 * IEmBackend::gpuRadioEnabled / downloadRadioWaveforms with BackendConfig::radio_
 * (corsika/kokkos/em/execution/IEmBackend.hpp:78,86; BackendConfig.hpp:66) and
 * the per-front RadioAccumulator::accumulateLeptonTracks call
 * (corsika/detail/kokkos/em/execution/LeptonCascade.inl:81-82): once attached,
 * every e+- step of every later lepton front (all pool slots) is projected onto
 * `radio` as a compact track (start/end position and time, charge-signed
 * weight). The projection runs asynchronously on the radio handle's stream and
 * overlaps the following fronts; c8mi_radio_download waits for it. NULL
 * detaches (after the projections in flight finish). The radio handle must not
 * be destroyed while attached. */
int c8mi_em_attach_radio(c8mi_em* em, c8mi_radio* radio);

/* ---- Per-shower radio of the pool ----
 * The reference gives every shower its own RadioAccumulator (BackendConfig::radio_,
 * corsika/kokkos/em/execution/BackendConfig.hpp:66) fed by each lepton front
 * (LeptonCascade.inl:81-82). Here every pool slot keeps its own observers,
 * int64 fixed-point waveforms and counters resident on the device, and one
 * projection per lepton front covers all slots, each slot's e+- steps on its
 * own observers. `config` gives the algorithms, refinement, exactness,
 * fixed-point limit and propagation (observer lists ignored; deterministic
 * only); each slot holds up to observers_per_slot (<= 256) observers of up
 * to bins_per_observer bins. Enabling (again) clears every slot's observers. */
int c8mi_em_enable_slot_radio(c8mi_em* em, const c8mi_radio_config* config, uint32_t observers_per_slot,
                              uint32_t bins_per_observer);
/* Sets a slot's observers (count 0: none) and clears its waveforms and
 * counters; applies to the lepton fronts that follow. c8mi_em_slot_begin (and
 * c8mi_em_begin_shower for slot 0) starts a shower without observers;
 * c8mi_em_serve sets every admitted request's own. */
int c8mi_em_slot_set_radio(c8mi_em* em, uint32_t slot, const c8mi_radio_observer* observers, uint64_t count);
/* RadioAccumulator::downloadFixed for one slot: [3][sum of its observers'
 * bins] per algorithm (either may be NULL) and its counters; waits. */
int c8mi_em_slot_download_radio(c8mi_em* em, uint32_t slot, int64_t* coreas, int64_t* zhs,
                                c8mi_radio_counters* counters);
/* Per-shower projection time (HIP events on its stream) and launches since
 * c8mi_em_enable_kernel_timing. */
int c8mi_em_slot_radio_kernel_time(c8mi_em* em, double* ms, uint64_t* launches);

/* Per-kernel device time (HIP events) since enable: ms[0] photon step,
 * ms[1] lepton transport, ms[2] lepton vertex, ms[3] count, ms[4] scan,
 * ms[5] scatter. */
int c8mi_em_enable_kernel_timing(c8mi_em* em, int enable);
int c8mi_em_kernel_time(const c8mi_em* em, double ms[6], uint64_t launches[6]);

#ifdef __cplusplus
}
#endif

#endif /* C8MI_EM_H */
