// extern "C" output sink entry points (include/c8mi/c8mi_output.h): the checks
// and writer layout of CorsikaOutputSink::onGpuProfile
// (corsika/detail/kokkos/output/CorsikaOutputSink.inl:117-197) and the
// energy_ledger block of ShowerReport
// (corsika/detail/kokkos/application/cascade/ShowerReport.inl:1221-1287).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>

#include "c8mi/c8mi_output.h"
#include "capi_guard.hpp"

using c8mi::guarded;

namespace {

double deposited_of(c8mi_em_ledger const& l) {
  return l.continuous_deposit + l.cut_deposit + l.native_muon_pair_cut_energy;
}

// The accumulated-output checks of onGpuProfile (CorsikaOutputSink.inl:130-156).
void check_ledger(c8mi_em_ledger const& l) {
  double const deposited = deposited_of(l);
  double const native = l.native_muon_pair_cut_energy;
  auto non_negative = [](double v) { return std::isfinite(v) && v >= 0.; };
  if (l.fixed_point_overflows != 0 || l.invalid_records != 0 || !std::isfinite(deposited) ||
      !std::isfinite(native) || native > 0. || deposited < native - 1.e-12 ||
      !non_negative(l.medium_rest_mass_input) || !non_negative(l.cut_rest_mass) ||
      !non_negative(l.observed_energy) || !non_negative(l.escaped_energy) ||
      !non_negative(l.photoelectric_binding) || !non_negative(l.observation_cut_overlap) ||
      !std::isfinite(l.mass_convention_correction))
    throw std::invalid_argument("output: the profile ledger contains invalid accumulated output");
}

}  // namespace

extern "C" {

int c8mi_em_output_profile(double const* photons, double const* electrons, double const* positrons,
                           double const* energy_loss, std::uint32_t bins, c8mi_em_ledger const* ledger,
                           double* longitudinal, double* energy_loss_em, double* energy_loss_muon,
                           c8mi_em_output_statistics* statistics) {
  return guarded([&] {
    if (!photons || !electrons || !positrons || !energy_loss || !ledger || !longitudinal || !energy_loss_em ||
        !energy_loss_muon || !statistics)
      throw std::invalid_argument("output: null argument");
    if (bins == 0) throw std::length_error("output: a profile needs at least one bin");
    check_ledger(*ledger);
    double const native = ledger->native_muon_pair_cut_energy;
    // every bin first: a rejected profile writes nothing
    for (std::uint32_t b = 0; b < bins; ++b) {
      double const v[3] = {photons[b], electrons[b], positrons[b]};
      for (double x : v)
        if (!std::isfinite(x) || x < 0.)
          throw std::invalid_argument("output: the profile contains a non-finite or negative bin");
      if (!std::isfinite(energy_loss[b]) || energy_loss[b] < native - 1.e-12)
        throw std::invalid_argument("output: the profile contains a non-finite or negative bin");
    }
    for (std::uint32_t b = 0; b < bins; ++b) {
      double* row = longitudinal + static_cast<std::uint64_t>(b) * C8MI_LONGITUDINAL_COLUMNS;
      row[C8MI_LONGITUDINAL_PHOTON] = photons[b];
      row[C8MI_LONGITUDINAL_ELECTRON] = electrons[b];
      row[C8MI_LONGITUDINAL_POSITRON] = positrons[b];
      row[C8MI_LONGITUDINAL_MUON_MINUS] = 0.;
      row[C8MI_LONGITUDINAL_MUON_PLUS] = 0.;
      // the EM share is the total less the muons' (none histogrammed here)
      energy_loss_em[b] = energy_loss[b];
      energy_loss_muon[b] = 0.;
    }
    statistics->steps = ledger->steps;
    statistics->deposited_steps = ledger->deposited_steps;
    statistics->weighted_deposited_energy_gev = deposited_of(*ledger);
    statistics->weighted_native_muon_pair_cut_energy_gev = native;
    statistics->weighted_muon_parent_productions = 0.;
  });
}

int c8mi_em_energy_ledger_report(c8mi_em_ledger const* ledgers, std::uint64_t count,
                                 c8mi_em_report_context const* context, c8mi_em_energy_ledger* out) {
  return guarded([&] {
    if ((count && !ledgers) || !context || !out) throw std::invalid_argument("report: null argument");
    if (!std::isfinite(context->initial_total_gev) || context->initial_total_gev < 0. ||
        !std::isfinite(context->thinning_fraction) || !std::isfinite(context->acceptance_tolerance) ||
        context->acceptance_tolerance < 0.)
      throw std::invalid_argument("report: invalid context");
    c8mi_em_energy_ledger r{};
    std::uint64_t fallbacks = 0;
    for (std::uint64_t k = 0; k < count; ++k) {
      c8mi_em_ledger const& l = ledgers[k];
      check_ledger(l);
      r.medium_rest_mass_input_gev += l.medium_rest_mass_input;
      r.deposited_gev += deposited_of(l);
      r.cut_rest_mass_energy_gev += l.cut_rest_mass;
      r.observed_total_energy_gev += l.observed_energy;
      r.escaped_total_energy_gev += l.escaped_energy;
      r.unwritten_photoelectric_binding_energy_gev += l.photoelectric_binding;
      r.observation_cut_overlap_energy_gev += l.observation_cut_overlap;
      r.mass_convention_correction_gev += l.mass_convention_correction;
      r.native_muon_pair_cut_energy_gev += l.native_muon_pair_cut_energy;
      r.fallback_energy_gev += l.fallback_energy;
      r.pending_energy_gev += l.pending_energy;
      r.thinning_discarded_energy_gev += l.thinning_discarded_energy;
      r.injected_gev += l.injected_energy;
      for (int reason = 0; reason < 32; ++reason) fallbacks += l.fallback_reasons[reason];
    }
    r.initial_total_gev = context->initial_total_gev;
    r.source_gev = r.initial_total_gev + r.medium_rest_mass_input_gev + r.mass_convention_correction_gev;
    r.terminal_gev = r.deposited_gev + r.cut_rest_mass_energy_gev + r.observed_total_energy_gev +
                     r.escaped_total_energy_gev + r.unwritten_photoelectric_binding_energy_gev -
                     r.observation_cut_overlap_energy_gev;
    r.residual_gev = r.source_gev - r.terminal_gev;
    r.extended_residual_gev =
        r.residual_gev - r.fallback_energy_gev - r.pending_energy_gev - r.thinning_discarded_energy_gev;
    r.relative_closure_error =
        std::fabs(r.residual_gev) / std::max(r.source_gev, std::numeric_limits<double>::min());
    r.acceptance_tolerance = context->acceptance_tolerance > 0. ? context->acceptance_tolerance : 1.e-4;
    r.scalar_em_steps = context->scalar_em_steps;
    r.fallbacks = fallbacks;
    r.complete_coverage = context->electromagnetic_primary != 0 && context->thinning_fraction <= 0. &&
                          context->scalar_em_steps == 0 && context->particles_returned_for_cpu_fallback == 0 &&
                          context->particles_returned_for_cpu_memory_spill == 0 &&
                          context->specified_cpu_final_states == 0 && fallbacks == 0 &&
                          r.pending_energy_gev == 0.;
    r.accepted = r.complete_coverage && r.relative_closure_error <= r.acceptance_tolerance;
    *out = r;
  });
}

}  // extern "C"
