/*
 * c8mi output sink: what the reference's CorsikaOutputSink and ShowerReport do
 * with a shower's device results once they are downloaded.
 *
 *  - c8mi_em_output_profile replaces CorsikaOutputSink::onGpuProfile
 *    (corsika/detail/kokkos/output/CorsikaOutputSink.inl:117-197): it checks the
 *    downloaded histograms and ledger as the sink does and lays them out as the
 *    CORSIKA writers take them (longitudinal particle columns, EM and muon
 *    energy loss per bin, the sink's step statistics).
 *  - c8mi_em_energy_ledger_report replaces the energy_ledger block of
 *    ShowerReport (corsika/detail/kokkos/application/cascade/ShowerReport.inl:
 *    1221-1287): source, terminal, residual and relative closure error by the
 *    reference's definition, its complete-coverage rule and its 1e-4
 *    acceptance, over the ledgers of every side of a shower (device and host).
 *
 * Both run on the host over plain arrays; neither touches the GPU.
 */
#ifndef C8MI_OUTPUT_H
#define C8MI_OUTPUT_H

#include <stdint.h>

#include "c8mi/c8mi_em.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Longitudinal columns of c8mi_em_output_profile, in this order per bin. */
enum {
  C8MI_LONGITUDINAL_PHOTON = 0,
  C8MI_LONGITUDINAL_ELECTRON = 1,
  C8MI_LONGITUDINAL_POSITRON = 2,
  C8MI_LONGITUDINAL_MUON_MINUS = 3,
  C8MI_LONGITUDINAL_MUON_PLUS = 4,
  C8MI_LONGITUDINAL_COLUMNS = 5
};

/* The sink's statistics of one profile (CorsikaOutputSink statistics_). */
typedef struct c8mi_em_output_statistics {
  uint64_t steps;
  uint64_t deposited_steps;
  double weighted_deposited_energy_gev;        /* continuous + cut deposits + native mu-pair cuts */
  double weighted_native_muon_pair_cut_energy_gev;
  double weighted_muon_parent_productions;     /* no muon production on this backend: 0 */
} c8mi_em_output_statistics;

/* Checks a downloaded profile (c8mi_em_download_profile: `bins` entries per
 * histogram, weights and GeV) and its ledger as onGpuProfile does, then writes
 * longitudinal[bins][C8MI_LONGITUDINAL_COLUMNS] (weighted particle crossings),
 * energy_loss_em[bins] and energy_loss_muon[bins] (GeV). This backend keeps no
 * muon histograms: the muon columns are zero, and the native mu-pair cut energy
 * (the muon energy loss's only device entry) is reported in the statistics.
 * Status 1 for an invalid profile (fixed-point overflow, invalid records,
 * non-finite or negative bins or ledger terms), 3 for zero bins. */
int c8mi_em_output_profile(const double* photons, const double* electrons, const double* positrons,
                           const double* energy_loss, uint32_t bins, const c8mi_em_ledger* ledger,
                           double* longitudinal, double* energy_loss_em, double* energy_loss_muon,
                           c8mi_em_output_statistics* statistics);

/* What the report needs beyond the ledgers (ShowerReport's beam, thinning and
 * router statistics). */
typedef struct c8mi_em_report_context {
  double initial_total_gev;               /* weighted total energy of the primaries */
  double thinning_fraction;               /* emthinfrac: > 0 with EM thinning on */
  double acceptance_tolerance;            /* 0: the reference's 1e-4 */
  uint32_t electromagnetic_primary;       /* is_em(beamCode) */
  uint32_t pad_;
  uint64_t scalar_em_steps;               /* host steps of photons and e+- */
  uint64_t particles_returned_for_cpu_fallback;
  uint64_t particles_returned_for_cpu_memory_spill;
  uint64_t specified_cpu_final_states;
} c8mi_em_report_context;

/* The energy_ledger block. The reference's definition: initial_total +
 * medium_rest_mass_input + mass_convention_correction (source) = deposited +
 * cut_rest_mass + observed_total + escaped_total + unwritten_photoelectric_binding
 * - observation_cut_overlap (terminal). The terms a complete-coverage shower
 * never has (fallbacks, pending particles, thinning) are reported beside it,
 * with the residual that books them as sinks. */
typedef struct c8mi_em_energy_ledger {
  double initial_total_gev;
  double medium_rest_mass_input_gev;
  double deposited_gev;
  double cut_rest_mass_energy_gev;
  double observed_total_energy_gev;
  double escaped_total_energy_gev;
  double unwritten_photoelectric_binding_energy_gev;
  double observation_cut_overlap_energy_gev;
  double mass_convention_correction_gev;
  double source_gev;
  double terminal_gev;
  double residual_gev;
  double relative_closure_error;
  double acceptance_tolerance;
  double native_muon_pair_cut_energy_gev;  /* inside deposited_gev */
  double fallback_energy_gev;
  double pending_energy_gev;
  double thinning_discarded_energy_gev;
  double extended_residual_gev;            /* residual less the three terms above */
  double injected_gev;                     /* the ledgers' injected energy, summed */
  uint64_t scalar_em_steps;
  uint64_t fallbacks;                      /* device and host fallback records, summed */
  uint32_t complete_coverage;
  uint32_t accepted;
} c8mi_em_energy_ledger;

/* Sums `count` ledgers (the sides of one shower) and fills the report. Status 1
 * for invalid ledgers (as c8mi_em_output_profile) or a null argument. */
int c8mi_em_energy_ledger_report(const c8mi_em_ledger* ledgers, uint64_t count,
                                 const c8mi_em_report_context* context, c8mi_em_energy_ledger* out);

#ifdef __cplusplus
}
#endif

#endif /* C8MI_OUTPUT_H */
