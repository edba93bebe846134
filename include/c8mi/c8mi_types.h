/*
 * c8mi — MI355X-native EM transport and radio for CORSIKA 8.
 *
 * Plain-old-data records that cross the C ABI. Each struct is layout-compatible
 * with the CORSIKA8-kokkos record it replaces, so a C++ adapter can pass the
 * reference's vectors by pointer without conversion (sizes are checked below).
 *
 *   c8mi_particle_state          <-> corsika::kokkos::ParticleState
 *                                    (corsika/kokkos/data/ParticleState.hpp:24-44, 128 B)
 *   c8mi_interaction_record      <-> corsika::kokkos::InteractionRecord
 *                                    (corsika/kokkos/em/InteractionRecord.hpp:48-93, 320 B)
 *   c8mi_lepton_transport_record <-> corsika::kokkos::LeptonTransportRecord
 *                                    (corsika/kokkos/em/LeptonTransportRecord.hpp:16-70, 752 B)
 *   c8mi_radio_observer          <-> corsika::kokkos::radio::RadioObserverSnapshot
 *                                    (corsika/kokkos/radio/RadioObserverSnapshot.hpp)
 *   c8mi_radio_counters          <-> corsika::kokkos::radio::RadioCounters
 *                                    (corsika/kokkos/radio/RadioProjectionData.hpp:74-95)
 *
 * Units follow the reference: GeV, metres, seconds, V/m (CoREAS), V s/m (ZHS).
 */
#ifndef C8MI_TYPES_H
#define C8MI_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* PDG identities used on the EM path (corsika/kokkos/em/ParticleId.hpp:17-23). */
enum {
  C8MI_PID_ELECTRON = 11,
  C8MI_PID_POSITRON = -11,
  C8MI_PID_MUON_MINUS = 13,
  C8MI_PID_MUON_PLUS = -13,
  C8MI_PID_PHOTON = 22
};

#if defined(__cplusplus)
#define C8MI_ALIGN16 alignas(16)
#else
#define C8MI_ALIGN16 _Alignas(16)
#endif

typedef struct c8mi_particle_state {
  C8MI_ALIGN16 int32_t pid;
  int32_t medium_id;
  uint32_t generation;
  uint32_t reserved;
  double energy_gev; /* total energy incl. rest mass */
  double position_m[3];
  double direction[3];
  double time_s;
  double weight;
  uint64_t history_id;
  uint64_t parent_history_id;
  uint64_t step_id;
  int32_t environment_node; /* -3 unassigned, -1 outside the tree */
  int32_t pad_;
  uint64_t pad2_;
} c8mi_particle_state;

typedef struct c8mi_interaction_record {
  c8mi_particle_state particle;
  int32_t deferred_photon_fallback_reason;
  int32_t deferred_photon_table_status;
  uint32_t interaction_vertex_reached;
  uint32_t pad0_;
  uint64_t input_index;
  int32_t process_id;
  int32_t status; /* InteractionStatus */
  uint64_t component_hash;
  double total_rate_cm2_per_g;
  double vertex_total_rate_cm2_per_g;
  double interaction_grammage_g_per_cm2;
  double mass_density_g_per_cm3;
  double particle_mass_gev;
  double decay_distance_m;
  double decay_uniform;
  uint64_t decay_draw_id;
  double energy_fraction;
  double distance_uniform;
  double process_uniform;
  double loss_quantile;
  uint64_t distance_draw_id;
  uint64_t process_draw_id;
  uint64_t loss_draw_id;
  double proposal_selection_uniform;
  uint32_t process_random_process_id;
  uint32_t proposal_selection_random_process_id;
  uint64_t proposal_selection_draw_id;
  uint64_t pad1_;
} c8mi_interaction_record;

typedef struct c8mi_lepton_transport_record {
  c8mi_interaction_record interaction;
  c8mi_particle_state start;
  c8mi_particle_state end;
  uint64_t input_index;
  int32_t limit; /* LeptonTransportLimit */
  int32_t start_layer_index;
  int32_t end_layer_index;
  int32_t pad0_;
  double distance_m;
  double traversed_grammage_g_per_cm2;
  double continuous_step_grammage_g_per_cm2;
  double start_density_g_per_cm3;
  double end_density_g_per_cm3;
  double limiting_radius_m;
  double continuous_deposited_energy_gev;
  double cut_deposited_energy_gev;
  uint32_t observation_surface_reached_before_cut;
  uint16_t multiple_scattering_applied;
  uint16_t multiple_scattering_status;
  uint32_t multiple_scattering_iterations;
  uint32_t pad1_;
  double multiple_scattering_angle_rad;
  double multiple_scattering_first_uniform;
  double multiple_scattering_second_uniform;
  double multiple_scattering_azimuth_uniform;
  uint32_t magnetic_bending_applied;
  uint32_t magnetic_step_status;
  double magnetic_step_limit_m;
  double magnetic_gyroradius_m;
  double magnetic_bend_parameter;
  double magnetic_chord_length_m;
} c8mi_lepton_transport_record;

/* Observer: position (m), window start/duration (s), sample rate (Hz), bins. */
typedef struct c8mi_radio_observer {
  double position_m[3];
  double start_time_s;
  double duration_s;
  double sample_rate_hz;
  uint64_t number_of_bins;
} c8mi_radio_observer;

/* Straight-ray propagation snapshot (FlatAtmosphereRadioSnapshot.hpp).
 * homogeneous_refractive_index > 0 selects n R / c and ignores the table.
 * Otherwise refractivity[i] = n(min_height + i*step) - 1 and integrated[] is
 * its running sum scaled by 1/step (see c8mi_radio_build_flat_table). */
typedef struct c8mi_radio_propagation {
  double homogeneous_refractive_index;
  double minimum_height_m;
  double maximum_height_m;
  double step_m;
  double inverse_step_per_m;
  double slope_refractivity_lower;
  double slope_integrated_refractivity_lower;
  double slope_refractivity_upper;
  double slope_integrated_refractivity_upper;
  const double* refractivity;
  const double* integrated_refractivity;
  uint64_t table_size;
} c8mi_radio_propagation;

/* RadioConfig (corsika/kokkos/radio/RadioConfig.hpp:21-53). */
typedef struct c8mi_radio_config {
  int32_t coreas_enabled;
  int32_t zhs_enabled;
  int32_t deterministic;             /* 1: int64 fixed point (production default) */
  uint32_t zhs_subtrack_refinement;  /* 1..64, default 1 */
  /* 1: evaluate in the reference's FP64 operation order, so the fixed-point
   * waveforms are bit-identical to its scalar projection. 0 (default): FP64
   * with hardware reciprocal / reciprocal-square-root refinement; field error
   * against the FP64 result below 1e-6 relative. */
  int32_t exact_fp64;
  int32_t reserved0;
  double fixed_point_field_limit_v_per_m; /* default 1 V/m */
  c8mi_radio_propagation propagation;
  const c8mi_radio_observer* coreas_observers;
  uint64_t coreas_observer_count;
  const c8mi_radio_observer* zhs_observers;
  uint64_t zhs_observer_count;
} c8mi_radio_config;

typedef struct c8mi_radio_counters {
  uint64_t coreas_contributions;
  uint64_t zhs_contributions;
  uint64_t zhs_subtracks;
  uint64_t valid_tracks;
  uint64_t fixed_point_overflows;
} c8mi_radio_counters;

/* Compact device-resident radio tracks, structure of arrays, FP64:
 * row k of a [9][stride] block holds field k for every track:
 *   0..2 start x,y,z (m)   3..5 end x,y,z (m)   6 start time (s)   7 end time (s)
 *   8 signed weight: -weight for e-, +weight for e+
 * 72 bytes per track. This is the native form the EM lepton front emits. */
enum { C8MI_RADIO_TRACK_FIELDS = 9 };

#if defined(__cplusplus)
static_assert(sizeof(c8mi_particle_state) == 128, "ParticleState layout");
static_assert(sizeof(c8mi_interaction_record) == 320, "InteractionRecord layout");
static_assert(sizeof(c8mi_lepton_transport_record) == 752, "LeptonTransportRecord layout");
#else
_Static_assert(sizeof(c8mi_particle_state) == 128, "ParticleState layout");
_Static_assert(sizeof(c8mi_interaction_record) == 320, "InteractionRecord layout");
_Static_assert(sizeof(c8mi_lepton_transport_record) == 752, "LeptonTransportRecord layout");
#endif

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* C8MI_TYPES_H */
