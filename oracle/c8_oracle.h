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

#ifdef __cplusplus
}
#endif

#endif /* C8_ORACLE_H */
