// Device EM physics of one particle step: interaction selection, photon and
// lepton transport, vertex reselection, final states, LPM suppression,
// thinning and the profile / ledger contributions of the step.
//
// Reference algorithms (corsika/detail/kokkos/em/...):
//   algorithms/InteractionSelection.inl:12-215, PhotonTransportStep.inl:57-261,
//   PhotonFinalStateStep.inl:61-825, LeptonContinuousStep.inl:43-159,
//   LeptonTransportStep.inl:27-513 (field-free form), LeptonVertexSelection.inl:41-199,
//   LeptonFinalStateStep.inl:131-1036, EmThinning.inl:28-104,
//   models/PhotonPairLpm.inl:58-157, models/BremsLpm.inl:58-164,
//   models/PhotonPairKinematics.inl:29-72,
//   tables/{DiscretePhysicsQueries,ContinuousPhysicsQueries,TableQueryResult}.inl,
//   corsika/detail/kokkos/output/algorithms/{ProfileAccumulation,ProfileProjection}.inl.
// Muons and the e+- pair (electron_pair) final state are handed to the host as
// fallback requests.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "em/em_device.hpp"
#include "em/environment.hpp"
#include "em/native_tables.hpp"
#include "em/philox.hpp"

namespace c8mi::em::phys {

// Working copy of a particle inside a step: the ABI state plus the Philox
// shower key of the shower it belongs to (children inherit it).
struct Particle : c8mi_particle_state {
  std::uint64_t shower_key;
};

// ---- constants (ScalarPhysicalConstants.hpp, TransportMass.hpp, PhotonPairKinematics.hpp,
//      ParticleId.hpp, ContinuousPhysicsQueries.hpp, LeptonContinuousStep.hpp) ----
inline constexpr double kC = 299792458.;
inline constexpr double kElementaryCharge = 1.6021766208e-19;
inline constexpr double kRigidity = 0.2997925146834389;
inline constexpr double kElectronMassGeV = 0.0005109989461;
inline constexpr double kTransportElectronMassGeV = 5.109989e-4;
inline constexpr double kTransportMuonMassGeV = 1.056584e-1;
inline constexpr double kCutTime = 10.e-3;
inline constexpr double kCutSafety = 0.9999;
inline constexpr double kMaxRelativeLoss = 0.10;
inline constexpr double kPhotonTwoPi = 6.283185307179586476925286766559;
inline constexpr double kPhotonPi = 3.1415926535897932384626433832795;
inline constexpr double kLeptonPi = 3.141592653589793238462643383279502884;
inline constexpr double kLeptonTwoPi = 6.283185307179586476925286766559005768;
inline constexpr double kHalfPrecision = 1.e-5;

enum Process : std::int32_t {
  kBrems = 1000000002,
  kIonization = 1000000003,
  kEpair = 1000000004,
  kCompton = 1000000010,
  kAnnihilation = 1000000012,
  kPhotonPair = 1000000013,
  kPhotonMuPair = 1000000015,
  kPhotoelectric = 1000000016,
  kPhotonPairFinalState = 2000000013,
  kPhotonMuPairFinalState = 2000000015
};
// Photon μ-pair draws (PhotonFinalStateRandomDomains.hpp:25-27).
inline constexpr std::uint64_t kMuPairSplitDraw = std::uint64_t{1} << 32;
inline constexpr std::uint64_t kMuPairAzimuthDraw = kMuPairSplitDraw + 1;
// reserved_ tag of a μ-pair muon below its transport mass (NativeMuonPairCut.hpp:17).
inline constexpr std::uint32_t kNativeMuonPairCutFlag = 1u;

enum Inter : int { kSelected = 0, kNoDiscrete = 1, kReselect = 2, kDistanceSampled = 3, kParticleCut = 4 };
enum Lookup : int {
  kTsSuccess = 0, kTsParticleNotFound = 1, kTsColumnNotFound = 2, kTsRateEnergy = 3, kTsInverseCdf = 4,
  kTsLossEnergy = 5, kTsLossQuantile = 6, kTsNonFinite = 7, kTsInvalidView = 9,
  kTsContParticle = 10, kTsContEnergy = 11, kTsContRange = 12, kTsTransportCut = 13
};
enum Reason : int {
  kFrUnsupportedParticle = 0, kFrUnsupportedGeometry = 2, kFrParticleTable = 3, kFrComponent = 4,
  kFrRateEnergy = 5, kFrInverseCdf = 6, kFrLossEnergy = 7, kFrLossQuantile = 8, kFrInvalidQuery = 9,
  kFrZeroRate = 11, kFrCpuOnly = 12, kFrNotImplemented = 13, kFrInvalidFinalState = 14,
  kFrLpmUnavailable = 15, kFrInvalidDensity = 16, kFrContEnergy = 17, kFrContRange = 18,
  kFrTransportCut = 19, kFrMoliereUnavailable = 20, kFrMoliereFailed = 21, kFrMagneticTransport = 22, kFrMagneticBoundary = 23, kFrGrammage = 24,
  kFrInverseGrammage = 25, kFrVertexLookup = 26, kFrSelectionReplay = 28
};
enum PhotonLimit : int { kPlInteraction = 0, kPlLayer = 1, kPlObservation = 2, kPlEscaped = 3, kPlCut = 4 };
enum LeptonLimit : int {
  kLlInteraction = 0, kLlContinuous = 1, kLlCut = 2, kLlLayer = 3, kLlObservation = 4,
  kLlEscaped = 5, kLlMagnetic = 6
};
enum Stage : std::uint32_t { kStageSelection = 1, kStageTransport = 2, kStageVertex = 3, kStageFinal = 4 };

C8MI_HD int sgn(double v) { return (0. < v) - (v < 0.); }

C8MI_HD double uniform(EmContext const& c, Particle const& p,
                       std::uint32_t process, std::uint64_t draw) {
  return uniform_open01(RandomKey{c.seed, p.shower_key, p.history_id, p.step_id, process, draw});
}

C8MI_HD double transport_mass(std::int32_t pid) {
  if (pid == 11 || pid == -11) return kTransportElectronMassGeV;
  if (pid == 13 || pid == -13) return kTransportMuonMassGeV;
  return 0.;
}
C8MI_HD double secondary_energy(std::int32_t pid, double total, double mass) {
  return (total - mass) + transport_mass(pid);
}
C8MI_HD double mass_correction(std::int32_t pid, double mass, double weight) {
  return weight * (transport_mass(pid) - mass);
}
C8MI_HD bool is_ee(std::int32_t pid) { return pid == 11 || pid == -11; }
C8MI_HD bool is_charged_lepton(std::int32_t pid) {
  return pid == 11 || pid == -11 || pid == 13 || pid == -13;
}
// is_native_muon_pair_cut (NativeMuonPairCut.inl:12-16): a muon of a photon
// μ-pair made at or below its transport mass, cut at its first step.
C8MI_HD bool native_muon_pair_cut(c8mi_particle_state const& p) {
  return p.reserved == kNativeMuonPairCutFlag && (p.pid == 13 || p.pid == -13) && p.generation > 0 &&
         p.step_id == 0 && p.energy_gev > 0. && p.energy_gev <= transport_mass(p.pid);
}

// ---- accumulation: the histograms and counters of the particle's shower ----
// hist / counters point at the block-private LDS copy of the shower (kept for
// the showers of the block's first and last particle) or at the shower's
// device totals.

struct Accumulator {
  long long* hist;      // 4 * bins
  long long* counters;  // kCounters
  EmContext const* ctx;
  ShowerSlot const* shower;
  std::uint32_t slot;
};

C8MI_HD void add_fixed(Accumulator const& a, long long* address, double value,
                       double scale) {
  double const scaled = value * scale;
  if (!isfinite(scaled) || fabs(scaled) >= 9.2233720368547748e18) {
    fetch_add(reinterpret_cast<unsigned long long*>(&a.counters[kOverflows]), 1ULL);
    return;
  }
  long long const inc = static_cast<long long>(rint(scaled));
  if (inc != 0) fetch_add(reinterpret_cast<unsigned long long*>(address), static_cast<unsigned long long>(inc));
}

C8MI_HD void count(Accumulator const& a, int index, unsigned long long n = 1) {
  fetch_add(reinterpret_cast<unsigned long long*>(&a.counters[index]), n);
}

C8MI_HD void ledger(Accumulator const& a, int term, double value) {
  add_fixed(a, &a.counters[term], value, a.ctx->energy_scale);
}

C8MI_HD double project_grammage(Accumulator const& a, double const p[3]) {
  ShowerSlot const& c = *a.shower;
  double proj = 0.;
  for (int k = 0; k < 3; ++k) proj += (p[k] - c.axis_start[k]) * c.axis_direction[k];
  double const fb = proj / c.axis_step;
  if (fb < 0.) return c.axis_grammage[0];
  if (!(fb < 1.8e19)) return c.axis_grammage[c.axis_support_count - 1];
  std::uint64_t const lower = static_cast<std::uint64_t>(fb);
  std::uint64_t const upper = lower + 1;
  if (upper >= c.axis_support_count) return c.axis_grammage[c.axis_support_count - 1];
  double const f = fb - static_cast<double>(lower);
  return c.axis_grammage[upper] * f + c.axis_grammage[lower] * (1. - f);
}

C8MI_HD void particle_profile(Accumulator const& a, std::int32_t pid, double start,
                     double end, double weight) {
  EmContext const& c = *a.ctx;
  double const s = start / c.bin_width, e = end / c.bin_width;
  if (c.crossing_mode == C8MI_PROFILE_CROSSING_FORWARD && !(e > s)) return;
  if (c.profile_bins == 0 || s == e || !(s >= -1.7976931348623157e308 && s <= 1.7976931348623157e308) ||
      !(e >= -1.7976931348623157e308 && e <= 1.7976931348623157e308))
    return;
  double const low = s < e ? s : e, high = s < e ? e : s;
  double first = floor(low) + 1., stop = floor(high) + 1.;
  double const limit = static_cast<double>(c.profile_bins);
  first = first < 0. ? 0. : first;
  stop = stop > limit ? limit : stop;
  if (!(first < stop)) return;
  int const h = pid == 22 ? kPhotonsHist : pid == 11 ? kElectronsHist : pid == -11 ? kPositronsHist : -1;
  if (h < 0) {
    count(a, kInvalid);
    return;
  }
  long long* hist = a.hist + static_cast<std::uint64_t>(h) * c.profile_bins;
  for (std::uint32_t k = static_cast<std::uint32_t>(first); k < static_cast<std::uint32_t>(stop); ++k)
    add_fixed(a, hist + k, weight, c.weight_scale);
}

C8MI_HD void energy_profile(Accumulator const& a, double start, double end, double deposit) {
  EmContext const& c = *a.ctx;
  if (!(deposit > 0.) || c.profile_bins == 0) return;
  count(a, kDepositedSteps);
  if (start > end) {
    double const t = start;
    start = end;
    end = t;
  }
  long long* hist = a.hist + static_cast<std::uint64_t>(kEnergyLossHist) * c.profile_bins;
  double const delta = end - start;
  int const maximum_bin = static_cast<int>(c.profile_bins - 1);
  double const w = c.bin_width;
  int first = static_cast<int>(start / w);
  first = first < 0 ? 0 : first;
  first = first > maximum_bin ? maximum_bin : first;
  if (delta < c.energy_loss_threshold) {
    add_fixed(a, hist + first, deposit, c.energy_scale);
    return;
  }
  int last = static_cast<int>(end / w);
  last = last < 0 ? 0 : last;
  last = last > maximum_bin ? maximum_bin : last;
  double const dens = deposit / delta;
  if (first == last) {
    add_fixed(a, hist + first, dens * delta, c.energy_scale);
    return;
  }
  add_fixed(a, hist + first, dens * (1. + static_cast<double>(first)) * w - dens * start, c.energy_scale);
  add_fixed(a, hist + last, dens * end - dens * static_cast<double>(last) * w, c.energy_scale);
  for (int bin = first + 1; bin < last; ++bin) add_fixed(a, hist + bin, dens * w, c.energy_scale);
}

C8MI_HD void point_energy_profile(Accumulator const& a, double x, double deposit) {
  EmContext const& c = *a.ctx;
  if (!(deposit > 0.) || c.profile_bins == 0) return;
  int const maximum_bin = static_cast<int>(c.profile_bins - 1);
  int bin = static_cast<int>(x / c.bin_width);
  bin = bin < 0 ? 0 : bin;
  bin = bin > maximum_bin ? maximum_bin : bin;
  count(a, kDepositedSteps);
  add_fixed(a, a.hist + static_cast<std::uint64_t>(kEnergyLossHist) * c.profile_bins + bin, deposit,
            c.energy_scale);
}

C8MI_HD void terminal_energy(Accumulator const& a, double e, bool observed) {
  if (!(e >= 0.) || !isfinite(e)) {
    count(a, kInvalid);
    return;
  }
  ledger(a, observed ? kObserved : kEscaped, e);
}

C8MI_HD void thinning_ledger(Accumulator const& a, int status, std::uint32_t multiplicity,
                    std::uint32_t kept, double parent_energy_weight,
                    double kept_energy_weight) {
  if (status == 1) count(a, kThinHillas);
  else if (status == 2) count(a, kThinStatistical);
  if (multiplicity > kept) count(a, kThinDiscarded, multiplicity - kept);
  ledger(a, kThinningEnergy, parent_energy_weight - kept_energy_weight);
}

// Observation / fallback records: appended to the shower's own region with one
// atomic per record (rare).
C8MI_HD void record_observation(Accumulator const& a, Particle const& p) {
  EmContext const& c = *a.ctx;
  unsigned long long const n = fetch_add(&c.record_counts[2u * a.slot], 1ULL);
  if (n < c.observation_capacity) c.observations[a.slot * c.observation_capacity + n] = p;
}

C8MI_HD void record_fallback(Accumulator const& a, Particle const& p, int reason,
                    std::int32_t process, std::uint32_t stage) {
  EmContext const& c = *a.ctx;
  count(a, kFallbackReason0 + (reason & 31));
  ledger(a, kFallbackEnergy, p.energy_gev * p.weight);
  unsigned long long const n = fetch_add(&c.record_counts[2u * a.slot + 1u], 1ULL);
  if (n < c.fallback_capacity) {
    c8mi_em_fallback f;
    f.particle = p;
    f.reason = reason;
    f.process_id = process;
    f.stage = stage;
    f.pad_ = 0;
    c.fallbacks[a.slot * c.fallback_capacity + n] = f;
  }
}

// The snapshot's photoelectric electron carries the history ID it is assigned
// (PhotonFinalStateStep.inl:649-660 sets it before the snapshot; every other
// process snapshots its children before their IDs): once the child's ID is
// known, the one child of the captured photoelectric vertex takes it.
C8MI_HD void patch_first_interaction(c8mi_em_first_interaction& r, unsigned long long candidates,
                                     c8mi_particle_state const& child) {
  if (candidates == 0 || child.generation != 1 || child.pid != 11 || r.process_id != kPhotoelectric ||
      r.secondary_count != 1 || child.parent_history_id != r.parent_at_vertex.history_id ||
      r.secondaries[0].history_id != r.parent_at_vertex.history_id)
    return;
  r.secondaries[0].history_id = child.history_id;
}

// The writers' first-interaction snapshot (PhotonFront.inl:95-111,
// LeptonFinalStateStep.inl:609-615): the first final state of a
// generation-zero parent, its children before thinning with the parent's
// weight. The first candidate of the shower writes it.
C8MI_HD void first_interaction(Accumulator const& a, Particle const& parent, std::int32_t process,
                               Particle const& first, Particle const* second) {
  EmContext const& c = *a.ctx;
  if (parent.generation != 0 || c.first_interactions == nullptr) return;
  if (fetch_add(&c.first_candidates[a.slot], 1ULL) != 0) return;
  c8mi_em_first_interaction& r = c.first_interactions[a.slot];
  r.parent_at_vertex = parent;
  r.secondaries[0] = first;
  r.secondaries[0].weight = parent.weight;
  c8mi_particle_state none{};
  r.secondaries[1] = second ? static_cast<c8mi_particle_state const&>(*second) : none;
  r.secondaries[1].weight = second ? parent.weight : 0.;
  r.secondaries[2] = none;
  r.process_id = process;
  r.secondary_count = second ? 2u : 1u;
}

// ---- table facades ----

C8MI_HD int from_native(std::uint32_t status, bool continuous) {
  switch (status) {
    case C8MI_QUERY_SUCCESS: return kTsSuccess;
    case C8MI_QUERY_PARTICLE_NOT_FOUND: return continuous ? kTsContParticle : kTsParticleNotFound;
    case C8MI_QUERY_COLUMN_NOT_FOUND: return kTsColumnNotFound;
    case C8MI_QUERY_ENERGY_OUT_OF_RANGE: return continuous ? kTsContEnergy : kTsRateEnergy;
    case C8MI_QUERY_QUANTILE_OUT_OF_RANGE: return kTsLossQuantile;
    case C8MI_QUERY_UNSUPPORTED_KINEMATICS:
    case C8MI_QUERY_ROOT_NOT_CONVERGED: return kTsInverseCdf;
    case C8MI_QUERY_NON_FINITE_INPUT: return kTsNonFinite;
    default: return kTsInvalidView;
  }
}

C8MI_HD int fallback_reason(int status) {
  switch (status) {
    case kTsParticleNotFound: return kFrParticleTable;
    case kTsColumnNotFound: return kFrComponent;
    case kTsRateEnergy: return kFrRateEnergy;
    case kTsInverseCdf: return kFrInverseCdf;
    case kTsLossEnergy: return kFrLossEnergy;
    case kTsLossQuantile: return kFrLossQuantile;
    case kTsContParticle: return kFrParticleTable;
    case kTsContEnergy: return kFrContEnergy;
    case kTsContRange: return kFrContRange;
    case kTsTransportCut: return kFrTransportCut;
    default: return kFrInvalidQuery;
  }
}

C8MI_HD bool requires_selected_loss(int reason) {
  return reason == kFrSelectionReplay || reason == kFrInverseCdf || reason == kFrLossEnergy ||
         reason == kFrLossQuantile;
}

struct Tq {
  int status;
  double value;
};

C8MI_HD Tq native(QueryResult const& r, bool continuous) {
  return {from_native(r.status, continuous), r.status == C8MI_QUERY_SUCCESS ? r.value : 0.};
}

C8MI_HD Tq cq_mass(EmContext const& c, std::int32_t pid) {
  auto const* u = detail::find_utility(c.tables, pid);
  if (!u) return {kTsContParticle, 0.};
  if (!isfinite(u->particle_mass_mev) || !(u->particle_mass_mev > 0.)) return {kTsInvalidView, 0.};
  return {kTsSuccess, u->particle_mass_mev};
}

C8MI_HD Tq cq_transport_cut(EmContext const& c, std::int32_t pid) {
  double const cut = (pid == 13 || pid == -13) ? c.muon_cut_mev : c.em_cut_mev;
  if (!isfinite(cut) || !(cut > 0.)) return {kTsInvalidView, 0.};
  return {kTsSuccess, cut};
}

C8MI_HD Tq cq_minimum_energy(EmContext const& c, std::int32_t pid) {
  auto const* u = detail::find_utility(c.tables, pid);
  if (!u) return {kTsContParticle, 0.};
  Tq const cut = cq_transport_cut(c, pid);
  if (cut.status != kTsSuccess) return cut;
  double const minimum = (1000. * transport_mass(pid) + cut.value) * kCutSafety;
  if (!isfinite(u->lower_energy_limit_mev) || u->lower_energy_limit_mev < u->particle_mass_mev ||
      !isfinite(minimum) || !(minimum > u->particle_mass_mev) || minimum < u->lower_energy_limit_mev ||
      minimum > u->spline.axis.high)
    return {kTsInvalidView, 0.};
  return {kTsSuccess, minimum};
}

C8MI_HD Tq cq_energy_after_loss(EmContext const& c, std::int32_t pid, double e0, double x) {
  if (!isfinite(x) || x < 0.) return {kTsNonFinite, 0.};
  Tq const initial = native(range(c.tables, pid, e0), true);
  if (initial.status != kTsSuccess) return initial;
  if (x >= initial.value && x != 0.) return {kTsTransportCut, 0.};
  return native(energy_after_loss(c.tables, pid, e0, x), true);
}

C8MI_HD bool pk_finite(double v) {
  return v == v && v <= 1.7976931348623157e308 && v >= -1.7976931348623157e308;
}

C8MI_HD bool decode_pair_split(double e, double s, double& rho) {
  if (!pk_finite(e) || !pk_finite(s) || !(e > 2. * kElectronMassGeV)) return false;
  double const tol = 64. * 2.2204460492503131e-16;
  if (s < -tol || s > 1. + tol) return false;
  double const b = s < 0. ? 0. : (s > 1. ? 1. : s);
  double const lower = kElectronMassGeV / e;
  double const width = 1. - 2. * lower;
  rho = lower + width * b;
  return pk_finite(rho) && rho >= lower && rho <= 1. - lower;
}

C8MI_HD bool encode_pair_split(double e, double rho, double& s) {
  if (!pk_finite(e) || !pk_finite(rho) || !(e > 2. * kElectronMassGeV)) return false;
  double const lower = kElectronMassGeV / e;
  double const width = 1. - 2. * lower;
  double const upper = 1. - lower;
  double const tol = 64. * 2.2204460492503131e-16;
  if (!(width > 0.) || rho < lower - tol || rho > upper + tol) return false;
  double const b = rho < lower ? lower : (rho > upper ? upper : rho);
  s = (b - lower) / width;
  return pk_finite(s) && s >= 0. && s <= 1.;
}

C8MI_HD Tq dq_loss_fraction(EmContext const& c, std::int32_t pid, std::int32_t process,
                   std::uint64_t comp, double e, double q) {
  if (process == kPhotonMuPairFinalState && pid == 22)
    return native(loss_fraction(c.tables, pid, kPhotonMuPair, comp, e, q, true), false);
  if (process == kPhotonPairFinalState && pid == 22) {
    Tq r = native(loss_fraction(c.tables, pid, kPhotonPair, comp, e, q, true), false);
    if (r.status != kTsSuccess) return r;
    double s = 0.;
    if (!encode_pair_split(e / 1000., r.value, s)) return {kTsInverseCdf, 0.};
    return {kTsSuccess, s};
  }
  return native(loss_fraction(c.tables, pid, process, comp, e, q, false), false);
}

struct Selection {
  int status;
  std::int32_t process;
  std::uint64_t component;
  double quantile;
  bool selected;
};

C8MI_HD Selection select_column(EmContext const& c, std::int32_t pid, double e, double u) {
  Selection r{kTsNonFinite, 0, 0, 0., false};
  if (!isfinite(u) || u < 0. || u >= 1.) return r;
  Tq const total = native(selection_total_rate(c.tables, pid, e), false);
  if (total.status != kTsSuccess) {
    r.status = total.status;
    return r;
  }
  r.status = kTsSuccess;
  if (!(total.value > 0.)) return r;
  double sampled = u * total.value;
  std::uint32_t begin, end;
  detail::dndx_range(c.tables, pid, begin, end);
  for (std::uint32_t pos = begin; pos < end; ++pos) {
    auto const* col = detail::selection_dndx(c.tables, pos);
    if (!col || col->pdg_id != pid) {
      r.status = kTsInvalidView;
      return r;
    }
    QueryResult rate{C8MI_QUERY_SUCCESS, 0, 0, 0.};
    if (!detail::rate_below_threshold(*col, e)) rate = rate_for_column(c.tables, *col, e);
    Tq const rt = native(rate, false);
    if (rt.status != kTsSuccess) {
      r.status = rt.status;
      r.process = col->process_id;
      r.component = col->component_hash;
      return r;
    }
    sampled -= rt.value;
    if (sampled < 0.) {
      double const quantile = (-sampled) / rt.value;
      r.process = col->process_id;
      r.component = col->component_hash;
      if (!isfinite(quantile) || quantile < 0. || quantile > 1.) {
        r.status = kTsInvalidView;
        return r;
      }
      r.quantile = quantile;
      r.selected = true;
      return r;
    }
  }
  return r;
}

enum Capability : int { kCapPair, kCapMuPair, kCapBrems, kCapCompton, kCapPhotoelectric,
                        kCapAnnihilation, kCapIonization, kCapCpuOnly, kCapNotImplemented };

C8MI_HD int capability(std::int32_t pid, std::int32_t process) {
  if (pid == 22 && process == kPhotonMuPair) return kCapMuPair;
  if (pid == 22 && process == kPhotonPair) return kCapPair;
  if (is_ee(pid) && process == kBrems) return kCapBrems;
  if (pid == 22 && process == kCompton) return kCapCompton;
  if (pid == 22 && process == kPhotoelectric) return kCapPhotoelectric;
  if (pid == -11 && process == kAnnihilation) return kCapAnnihilation;
  if (is_charged_lepton(pid) && process == kIonization) return kCapIonization;
  if (is_charged_lepton(pid) && process == kEpair) return kCapCpuOnly;
  if ((pid == 13 || pid == -13) && process == kBrems) return kCapBrems;
  switch (process) {
    case 1000000005: case 1000000006: case 1000000007: case 1000000008: case 1000000011:
    case 1000000014:
      return kCapCpuOnly;
    default: return kCapNotImplemented;
  }
}

C8MI_HD bool requires_replay(std::int32_t pid, std::int32_t process, double q) {
  int const cap = capability(pid, process);
  if (cap == kCapCpuOnly) return true;
  if (cap == kCapMuPair || cap == kCapNotImplemented) return false;
  return q < 0. || q > 1.;
}

// ---- interaction record ----

struct Interaction {
  Particle particle;
  std::int32_t deferred_reason;
  std::int32_t status;
  std::int32_t process_id;
  std::uint32_t vertex_reached;
  std::uint64_t component_hash;
  double total_rate;
  double interaction_grammage;
  double mass_density;
  double particle_mass_gev;
  double energy_fraction;
  double process_uniform;
  double loss_quantile;
};

struct Fallback {
  bool flag;
  int reason;
};

C8MI_HD void select_interaction(EmContext const& c, Particle const& p, Interaction& in,
                       Fallback& fb) {
  in.particle = p;
  in.deferred_reason = -1;
  in.status = kSelected;
  in.process_id = 0;
  in.vertex_reached = 0;
  in.component_hash = 0;
  in.total_rate = 0.;
  in.interaction_grammage = 0.;
  in.mass_density = 0.;
  in.particle_mass_gev = 0.;
  in.energy_fraction = 0.;
  in.process_uniform = 0.;
  in.loss_quantile = 0.;
  double const e_mev = p.energy_gev * 1000.;
  if (native_muon_pair_cut(p) || p.time_s > kCutTime) {
    in.status = kParticleCut;
    in.interaction_grammage = __builtin_inf();
    return;
  }
  bool const charged = is_charged_lepton(p.pid);
  if (p.pid == 22 && e_mev < c.em_cut_mev) {
    in.status = kParticleCut;
    in.interaction_grammage = __builtin_inf();
    return;
  }
  if (charged) {
    Tq const mass = cq_mass(c, p.pid);
    if (mass.status == kTsSuccess) in.particle_mass_gev = mass.value / 1000.;
    double const tm = 1000. * transport_mass(p.pid);
    bool const tm_ok = mass.status == kTsSuccess && tm > 0.;
    Tq const minimum = cq_minimum_energy(c, p.pid);
    Tq const cut = cq_transport_cut(c, p.pid);
    bool const below = tm_ok && cut.status == kTsSuccess && e_mev - tm < cut.value;
    if (below || (minimum.status == kTsSuccess && e_mev <= minimum.value)) {
      in.status = kParticleCut;
      in.interaction_grammage = __builtin_inf();
      return;
    }
  }
  Tq const total = native(total_rate(c.tables, p.pid, e_mev), false);
  if (total.status != kTsSuccess) {
    fb = {true, fallback_reason(total.status)};
    return;
  }
  in.total_rate = total.value;
  if (!(total.value > 0.)) {
    in.status = kNoDiscrete;
    in.interaction_grammage = __builtin_inf();
    return;
  }
  double const ud = uniform(c, p, kDistanceProcess, 0);
  in.interaction_grammage = -log(ud) / total.value;
  in.process_uniform = uniform(c, p, kColumnProcess, 0);
  if (charged) {
    in.status = kDistanceSampled;
    return;
  }
  double const us = uniform(c, p, kSelectionProcess, 0);
  Selection const s = select_column(c, p.pid, e_mev, us);
  if (s.status != kTsSuccess) {
    fb = {true, fallback_reason(s.status)};
    return;
  }
  if (!s.selected) {
    fb = {true, kFrZeroRate};
    return;
  }
  in.process_id = s.process;
  in.component_hash = s.component;
  in.loss_quantile = s.quantile;
  if (requires_replay(p.pid, s.process, s.quantile)) {
    in.deferred_reason = kFrSelectionReplay;
    return;
  }
  Tq const loss = dq_loss_fraction(c, p.pid, s.process, s.component, e_mev, s.quantile);
  if (loss.status != kTsSuccess) {
    int const reason = fallback_reason(loss.status);
    if (requires_selected_loss(reason)) {
      in.deferred_reason = reason;
      in.energy_fraction = loss.value;
      return;
    }
    fb = {true, reason};
    return;
  }
  in.energy_fraction = loss.value;
}

// ---- final-state helpers ----

C8MI_HD double bounded_cosine(double v) {
  double const m = v < 1. ? v : 1.;
  return -1. > m ? -1. : m;
}

C8MI_HD double sauter_cosine(double e, double u) {
  double const m2 = kElectronMassGeV * kElectronMassGeV;
  double const t = e * e - m2;
  double const p = sqrt(0. > t ? 0. : t);
  double const coord = 2. * u - 1.;
  double const emp = m2 / (e + p);
  double const omc = emp * (1. - coord) / (p * coord + e);
  return bounded_cosine(1. - omc);
}

C8MI_HD bool compton_cosines(double e, double v, double& pc, double& ec) {
  double const scattered = e * (1. - v);
  double const pe2 = 2. * v * e * kElectronMassGeV + v * v * e * e;
  if (!isfinite(e) || !(e > 0.) || !isfinite(v) || !(v > 0.) || !(v < 1.) || !(scattered > 0.) ||
      !(pe2 > 0.))
    return false;
  pc = 1. - (v * kElectronMassGeV) / scattered;
  ec = v * (e + kElectronMassGeV) / sqrt(pe2);
  double const tol = 1.e-12;
  if (!isfinite(pc) || !isfinite(ec) || pc < -1. - tol || pc > 1. + tol || ec < -1. - tol || ec > 1. + tol)
    return false;
  pc = bounded_cosine(pc);
  ec = bounded_cosine(ec);
  return true;
}

C8MI_HD bool photoelectric_energy(c8mi_photon_pair_lpm const& s, std::uint64_t comp,
                         double e, double& total, double& fraction) {
  if (!isfinite(e) || !(e > 0.) || s.component_count == 0 || s.component_count > 16 ||
      !isfinite(s.fine_structure_constant) || !(s.fine_structure_constant > 0.))
    return false;
  double z = 0.;
  for (std::uint32_t i = 0; i < s.component_count; ++i)
    if (s.components[i].component_hash == comp) {
      z = s.components[i].nuclear_charge;
      break;
    }
  if (!isfinite(z) || !(z > 0.)) return false;
  double const za = z * s.fine_structure_constant;
  double const binding = za * za * kElectronMassGeV / 2.;
  double const kinetic = e - binding;
  if (!isfinite(binding) || !isfinite(kinetic) || !(kinetic >= 0.)) return false;
  total = kElectronMassGeV + kinetic;
  fraction = kinetic / e;
  return isfinite(total) && total >= kElectronMassGeV && isfinite(fraction) && fraction >= 0. &&
         fraction <= 1.;
}

C8MI_HD void deflect(double const in[3], double cosine, double azimuth, double out[3],
            bool bound) {
  double const t = in[0] * in[0] + in[1] * in[1];
  double const tr = sqrt(0. > t ? 0. : t);
  double cp = 1., sp = 0.;
  if (tr > 0.) {
    cp = in[0] / tr;
    sp = in[1] / tr;
  }
  double const ct = in[2], st = tr;
  double const rx[3] = {ct * cp, ct * sp, -st};
  double const ry[3] = {-sp, cp, 0.};
  double c = cosine;
  if (bound) {
    double const m = 1. < c ? 1. : c;
    c = -1. > m ? -1. : m;
  }
  double const s2 = (1. - c) * (1. + c);
  double const s = sqrt(0. > s2 ? 0. : s2);
  double const lx = s * cos(azimuth), ly = s * sin(azimuth);
  double const z2 = 1. - lx * lx - ly * ly;
  double lz = sqrt(0. > z2 ? 0. : z2);
  if (c < 0.) lz = -lz;
  for (int a = 0; a < 3; ++a) out[a] = lz * in[a] + lx * rx[a] + ly * ry[a];
}

struct Thin {
  int status;  // 0 not applied, 1 Hillas, 2 statistical, 3 invalid
  std::uint32_t keep;
  double w1, w2;
};

C8MI_HD Thin em_thinning(c8mi_em_thinning const& c, bool enabled, double pe, double pw,
                double e1, double e2, double u1, double u2) {
  if (!isfinite(pe) || !(pe > 0.) || !isfinite(pw) || pw < 0. || !isfinite(e1) || !(e1 > 0.) ||
      !isfinite(e2) || !(e2 > 0.) || !isfinite(u1) || !(u1 > 0.) || !(u1 < 1.) || !isfinite(u2) ||
      !(u2 > 0.) || !(u2 < 1.))
    return {3, 0U, 0., 0.};
  if (!enabled) return {0, 3U, pw, pw};
  if (!isfinite(c.threshold_gev) || !(c.threshold_gev > 0.) || !isfinite(c.maximum_weight) ||
      !(c.maximum_weight > 0.))
    return {3, 0U, 0., 0.};
  if (pw >= c.maximum_weight || pe > c.threshold_gev) return {0, 3U, pw, pw};
  double const sum = e1 + e2;
  if (!isfinite(sum) || !(sum > 0.)) return {3, 0U, 0., 0.};
  double const p1 = e1 / sum, p2 = e2 / sum;
  double const f1 = 1. / p1, f2 = 1. / p2, fm = c.maximum_weight / pw;
  Thin o{0, 0U, 0., 0.};
  if (f1 <= fm && f2 <= fm) {
    o.status = 1;
    if (u1 <= p1) {
      o.keep = 1U;
      o.w1 = pw * f1;
    } else {
      o.keep = 2U;
      o.w2 = pw * f2;
    }
  } else {
    o.status = 2;
    double const l1 = fmin(f1, fm), l2 = fmin(f2, fm);
    if (u1 * l1 <= 1.) {
      o.keep |= 1U;
      o.w1 = pw * l1;
    }
    if (u2 * l2 <= 1.) {
      o.keep |= 2U;
      o.w2 = pw * l2;
    }
  }
  if (c.erase_zero_weight == 0) {
    o.keep = 3U;
  } else {
    if (o.w1 == 0.) o.keep &= ~1U;
    if (o.w2 == 0.) o.keep &= ~2U;
  }
  return o;
}

C8MI_HD bool two_child_thinning(EmContext const& c, Particle const& parent,
                       std::int32_t process, double e1, double e2, bool lepton,
                       Thin& t) {
  double const u1 = uniform(c, parent, static_cast<std::uint32_t>(process), 0x100);
  double const u2 = uniform(c, parent, static_cast<std::uint32_t>(process), 0x101);
  bool const enabled = c.thinning.enabled != 0 && !(lepton && !is_ee(parent.pid));
  t = em_thinning(c.thinning, enabled, parent.energy_gev, parent.weight, e1, e2, u1, u2);
  return t.status != 3;
}

// LPM tails shared by the photon-pair and bremsstrahlung suppressions.
C8MI_HD void lpm_functions(double sp, double s1, double gamma, double& xi, double& fi,
                  double& g) {
  constexpr double Pi = 3.141592653589793238462643383279502884;
  double const h = log(sp) / log(s1);
  xi = 1.;
  if (sp < s1) xi = 2.;
  else if (sp < 1.) xi = 1. + h - 0.08 * (1. - h) * (1. - (1. - h) * (1. - h)) / log(s1);
  double const s = sp / sqrt(xi) * gamma;
  double const s2 = s * s;
  if (s < 1.54954)
    fi = 1. - exp(-6. * s * (1. + (3. - Pi) * s) + s2 * s / (0.623 + 0.796 * s + 0.658 * s2));
  else
    fi = 1. - 0.012 / (s2 * s2);
  if (s < 0.710390) {
    double const psi =
        1. - exp(-4. * s - 8. * s2 / (1. + 3.936 * s + 4.97 * s2 - 0.05 * s2 * s + 7.50 * s2 * s2));
    g = 3. * psi - 2. * fi;
  } else if (s < 0.904912) {
    g = 36. * s2 / (36. * s2 + 1.);
  } else {
    g = 1. - 0.022 / (s2 * s2);
  }
}

C8MI_HD int pair_lpm(c8mi_photon_pair_lpm const& s, std::uint64_t comp, double e, double x,
            double rho, double& prob) {
  if (s.component_count == 0 || s.component_count > 16 || !isfinite(s.baseline_mass_density_g_per_cm3) ||
      !(s.baseline_mass_density_g_per_cm3 > 0.) || !isfinite(s.molecular_density_per_cm3) ||
      !(s.molecular_density_per_cm3 > 0.) || !isfinite(s.sum_charge) || !(s.sum_charge > 0.) ||
      !isfinite(s.e_lpm_mev) || !(s.e_lpm_mev > 0.) || !isfinite(s.classical_electron_radius_cm) ||
      !(s.classical_electron_radius_cm > 0.) || !isfinite(s.fine_structure_constant) ||
      !(s.fine_structure_constant > 0.))
    return 2;
  if (!isfinite(e) || !(e > 0.) || !isfinite(x) || !(x > 0.) || !(x < 1.) || !isfinite(rho) || !(rho > 0.))
    return 3;
  int found = -1;
  for (std::uint32_t i = 0; i < s.component_count; ++i)
    if (s.components[i].component_hash == comp) {
      found = static_cast<int>(i);
      break;
    }
  if (found < 0) return 1;
  double const z = s.components[found].nuclear_charge, rl = s.components[found].radiation_log_constant;
  if (!isfinite(z) || !(z > 0.) || !isfinite(rl) || !(rl > 0.)) return 2;
  constexpr double Pi = 3.141592653589793238462643383279502884;
  constexpr double SqrtTwo = 1.414213562373095048801688724209698079;
  double const dc = rho / s.baseline_mass_density_g_per_cm3;
  double const zm = pow(z, -1. / 3.);
  double s1 = 1. / (zm * rl);
  s1 = s1 * s1 * SqrtTwo;
  double const sp = 0.125 * sqrt(s.e_lpm_mev / (dc * e * x * (1. - x)));
  double const g0 = s.classical_electron_radius_cm / (s.fine_structure_constant * x);
  double const gamma = 1. + 4. * Pi * s.sum_charge * s.classical_electron_radius_cm * g0 * g0 *
                                s.molecular_density_per_cm3 * dc;
  double xi, fi, g;
  lpm_functions(sp, s1, gamma, xi, fi, g);
  double const p = ((xi / 3.) * (g / (gamma * gamma) + 2. * (x * x + (1. - x) * (1. - x)) * fi / gamma)) /
                   (1. - (4. / 3.) * x * (1. - x));
  if (!isfinite(p) || p < 0.) return 4;
  prob = p;
  return 0;
}

C8MI_HD int brems_lpm(c8mi_brems_lpm const& s, std::uint64_t comp, double e, double v,
             double rho, double& prob) {
  if (s.component_count == 0 || s.component_count > 16 || !isfinite(s.baseline_mass_density_g_per_cm3) ||
      !(s.baseline_mass_density_g_per_cm3 > 0.) || !isfinite(s.molecular_density_per_cm3) ||
      !(s.molecular_density_per_cm3 > 0.) || !isfinite(s.sum_charge) || !(s.sum_charge > 0.) ||
      !isfinite(s.e_lpm_mev) || !(s.e_lpm_mev > 0.) || !isfinite(s.lepton_mass_mev) ||
      !(s.lepton_mass_mev > 0.) || !isfinite(s.electron_mass_mev) || !(s.electron_mass_mev > 0.) ||
      !isfinite(s.muon_mass_mev) || !(s.muon_mass_mev > 0.) || !isfinite(s.classical_electron_radius_cm) ||
      !(s.classical_electron_radius_cm > 0.) || !isfinite(s.fine_structure_constant) ||
      !(s.fine_structure_constant > 0.))
    return 2;
  if (!isfinite(e) || !(e > s.lepton_mass_mev) || !isfinite(v) || !(v > 0.) || !(v < 1.) ||
      !isfinite(rho) || !(rho > 0.))
    return 3;
  int found = -1;
  for (std::uint32_t i = 0; i < s.component_count; ++i)
    if (s.components[i].component_hash == comp) {
      found = static_cast<int>(i);
      break;
    }
  if (found < 0) return 1;
  auto const& cp = s.components[found];
  if (!isfinite(cp.nuclear_charge) || !(cp.nuclear_charge > 0.) || !isfinite(cp.atomic_mass_number) ||
      !(cp.atomic_mass_number > 0.) || !isfinite(cp.radiation_log_constant) ||
      !(cp.radiation_log_constant > 0.))
    return 2;
  constexpr double Pi = 3.141592653589793238462643383279502884;
  constexpr double SqrtTwo = 1.414213562373095048801688724209698079;
  double const dc = rho / s.baseline_mass_density_g_per_cm3;
  double const zm = pow(cp.nuclear_charge, -1. / 3.);
  double const dn = 1.54 * pow(cp.atomic_mass_number, 0.27);
  double const na = s.lepton_mass_mev / s.muon_mass_mev * dn;
  double s1 = s.electron_mass_mev / (s.lepton_mass_mev * zm * cp.radiation_log_constant);
  s1 = s1 * s1 * (1. + na * na) * SqrtTwo;
  double const sp = 0.125 * sqrt(s.e_lpm_mev * v / (dc * e * (1. - v)));
  double const g0 = s.classical_electron_radius_cm * s.electron_mass_mev /
                    (s.fine_structure_constant * s.lepton_mass_mev * v);
  double const gamma = 1. + 4. * Pi * s.sum_charge * s.classical_electron_radius_cm * g0 * g0 *
                                s.molecular_density_per_cm3 * dc;
  double xi, fi, g;
  lpm_functions(sp, s1, gamma, xi, fi, g);
  double const p = ((xi / 3.) * (v * v * g / (gamma * gamma) + 2. * (1. + (1. - v) * (1. - v)) * fi / gamma)) /
                   ((4. / 3.) * (1. - v) + v * v);
  if (!isfinite(p) || p < 0.) return 4;
  prob = p;
  return 0;
}

C8MI_HD double annihilation_term(double a1, double a2, double rho) {
  return a1 * log(rho) + a2 / rho - rho;
}

C8MI_HD bool annihilation_rho(double e, double m, double u, double& rho) {
  double const gamma = e / m;
  if (!isfinite(gamma) || !(gamma > 1.) || !isfinite(u) || !(u > 0.) || !(u < 1.)) return false;
  double const aux = sqrt((gamma - 1.) / (gamma + 1.));
  double lower = 0.5 * (1. - aux), upper = 0.5 * (1. + aux);
  double const a2 = 1. / ((gamma + 1.) * (gamma + 1.));
  double const a1 = 1. + 2. * gamma * a2;
  double const lt = annihilation_term(a1, a2, lower), ut = annihilation_term(a1, a2, upper);
  double lv = annihilation_term(a1, a2, lower) - lt - u * (ut - lt);
  double const uv = annihilation_term(a1, a2, upper) - lt - u * (ut - lt);
  if (!isfinite(lv) || !isfinite(uv) || lv * uv > 0.) return false;
  double const precision = lower * (e * 1000.) * kHalfPrecision;
  if (!isfinite(precision) || !(precision > 0.)) return false;
  for (int it = 0; it <= 100; ++it) {
    double const center = (lower + upper) / 2.;
    double const cv = annihilation_term(a1, a2, center) - lt - u * (ut - lt);
    if (sgn(cv) == sgn(lv)) {
      lower = center;
      lv = cv;
    } else {
      upper = center;
    }
    if (fabs(upper - lower) < precision) break;
  }
  rho = lower;
  return isfinite(rho) && rho > 0. && rho < 1.;
}

C8MI_HD Particle child_of(Particle const& parent, std::int32_t pid, double energy,
                          double weight) {
  Particle c = parent;
  c.pid = pid;
  c.energy_gev = energy;
  c.parent_history_id = parent.history_id;
  c.generation = parent.generation + 1;
  c.step_id = 0;
  c.reserved = 0;
  c.weight = weight;
  return c;
}

// ---- step outputs ----

// Outputs of one source: up to two particles, stored as soon as they are made
// into the source's two staging entries (so they do not stay live in
// registers through the rest of the step).
struct Outputs {
  c8mi_particle_state* staging;
  double* track;               // radio track of the step: [9] rows `track_stride` apart; null: none
  std::uint64_t track_stride;
  std::uint32_t kind[2];  // SlotKind
  std::uint32_t next, other, children;
  bool step, interaction;
  bool tracked;  // `track` holds this step's e+- track
};

// Compact radio track of one e+- step (the make_radio_track_kinematics input,
// RadioProjection.inl:145-168): start and end positions, start and end times,
// weight signed by charge (negative for e-).
C8MI_HD void write_track(double* t, std::uint64_t stride, c8mi_particle_state const& start,
                         c8mi_particle_state const& end) {
  t[0] = start.position_m[0];
  t[stride] = start.position_m[1];
  t[2 * stride] = start.position_m[2];
  t[3 * stride] = end.position_m[0];
  t[4 * stride] = end.position_m[1];
  t[5 * stride] = end.position_m[2];
  t[6 * stride] = start.time_s;
  t[7 * stride] = end.time_s;
  t[8 * stride] = start.pid == 11 ? -start.weight : start.weight;
}

// A source without an e+- step record: a zero-length track, which the projection rejects.
C8MI_HD void write_no_track(double* t, std::uint64_t stride) {
  for (int k = 0; k < 9; ++k) t[k * stride] = 0.;
}

C8MI_HD void emit(Outputs& o, Particle const& p, std::uint32_t kind) {
  std::uint32_t const k = (o.kind[0] == kSlotNone) ? 0u : 1u;
  o.staging[k] = p;
  o.kind[k] = kind;
  if (kind == kSlotOtherChild) ++o.other;
  else ++o.next;
  if (kind != kSlotNextContinue) ++o.children;
}

}  // namespace c8mi::em::phys
