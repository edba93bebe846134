// TEST INFRASTRUCTURE ONLY: a C-ABI probe over the reference's own header-only
// functions, compiled in place from /root/reference by oracle/ref.mk into
// oracle/_ref/libref.so. It lets the tests check this repository's C
// restatement (and the HIP kernels) against the reference code itself on the
// same inputs. Only headers that compile unmodified are included; nothing of
// the reference is copied here.
#include <corsika/kokkos/em/algorithms/EmThinning.hpp>
#include <corsika/kokkos/em/models/BremsLpm.hpp>
#include <corsika/kokkos/em/models/MoliereScattering.hpp>
#include <corsika/kokkos/em/models/PhotonPairKinematics.hpp>
#include <corsika/kokkos/em/models/PhotonPairLpm.hpp>
#include <corsika/kokkos/em/random/Philox.hpp>
#include <corsika/kokkos/em/tables/MoliereInterpolation.hpp>
#include <corsika/kokkos/em/tables/ProposalNativeQueries.hpp>
#include <corsika/kokkos/em/tables/ProposalNativeTable.hpp>
#include <corsika/kokkos/utility/Sha256.hpp>
#include <corsika/kokkos/environment/EnvironmentBorrowedView.hpp>
#include <corsika/kokkos/environment/EnvironmentQuery.hpp>
#include <corsika/kokkos/environment/ObservationPlane.hpp>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "c8mi/c8mi_em.h"
#include "c8mi/c8mi_tables.h"

namespace ck = corsika::kokkos;

// Layout contract: c8mi's table records are byte-compatible with the
// reference's, so a reference export can be passed through unchanged.
static_assert(sizeof(c8mi_native_axis) == sizeof(ck::NativeAxisDescriptor));
static_assert(sizeof(c8mi_native_bicubic) == sizeof(ck::NativeBicubicDescriptor));
static_assert(sizeof(c8mi_native_cubic) == sizeof(ck::NativeCubicDescriptor));
static_assert(sizeof(c8mi_native_dndx_column) == sizeof(ck::NativeDndxColumn));
static_assert(offsetof(c8mi_native_dndx_column, spline) == offsetof(ck::NativeDndxColumn, spline_));
static_assert(offsetof(c8mi_native_dndx_column, relative_cut) ==
              offsetof(ck::NativeDndxColumn, relativeCut_));
static_assert(offsetof(c8mi_native_dndx_column, selection_ordinal) ==
              offsetof(ck::NativeDndxColumn, selectionOrdinal_));
static_assert(sizeof(c8mi_native_total_rate_column) == sizeof(ck::NativeTotalRateColumn));
static_assert(offsetof(c8mi_native_total_rate_column, lower_energy_limit_mev) ==
              offsetof(ck::NativeTotalRateColumn, lowerEnergyLimitMeV_));
static_assert(sizeof(c8mi_native_dedx_column) == sizeof(ck::NativeDedxColumn));
static_assert(offsetof(c8mi_native_dedx_column, accumulation_ordinal) ==
              offsetof(ck::NativeDedxColumn, accumulationOrdinal_));
static_assert(sizeof(c8mi_native_utility_column) == sizeof(ck::NativeUtilityColumn));
static_assert(offsetof(c8mi_native_utility_column, reverse) ==
              offsetof(ck::NativeUtilityColumn, reverse_));
static_assert(sizeof(c8mi_native_constants) == sizeof(ck::NativePhysicalConstants));
static_assert(sizeof(c8mi_native_result) == sizeof(ck::NativeQueryResult));
static_assert(sizeof(c8mi_photon_pair_lpm) == sizeof(ck::PhotonPairLpmSnapshot));
static_assert(sizeof(c8mi_brems_lpm) == sizeof(ck::BremsLpmSnapshot));
static_assert(offsetof(c8mi_brems_lpm, muon_e_lpm_mev) == offsetof(ck::BremsLpmSnapshot, muonELpmMeV_));
static_assert(sizeof(c8mi_em_thinning) == sizeof(ck::EmThinningConfig));

namespace {

template <class T>
struct Span {
  T const* pointer;
  std::size_t count;
  T const* data() const { return pointer; }
  std::size_t size() const { return count; }
};

template <class To, class From>
Span<To> span(From const* pointer, std::size_t count) {
  return {reinterpret_cast<To const*>(pointer), count};
}

ck::ProposalNativeDeviceView make_view(c8mi_native_table_set const& t) {
  using namespace ck;
  ck::NativePhysicalConstants constants{};
  std::memcpy(&constants, &t.constants, sizeof(constants));
  return ProposalNativeDeviceView(
      NativeArrayView<NativeDndxColumn, std::uint32_t>(
          span<NativeDndxColumn>(t.dndx_columns, t.dndx_column_count)),
      NativeArrayView<NativeTotalRateColumn, std::uint32_t>(
          span<NativeTotalRateColumn>(t.total_rate_columns, t.total_rate_column_count)),
      NativeArrayView<std::uint32_t, std::uint32_t>(
          span<std::uint32_t>(t.selection_column_indices, t.dndx_column_count)),
      NativeArrayView<NativeDedxColumn, std::uint32_t>(
          span<NativeDedxColumn>(t.dedx_columns, t.dedx_column_count)),
      NativeArrayView<std::uint32_t, std::uint32_t>(
          span<std::uint32_t>(t.dedx_accumulation_indices, t.dedx_column_count)),
      NativeArrayView<NativeUtilityColumn, std::uint32_t>(
          span<NativeUtilityColumn>(t.utility_columns, t.utility_column_count)),
      NativeArrayView<double, std::uint64_t>(span<double>(t.bicubic_values, t.bicubic_coefficient_count)),
      NativeArrayView<double, std::uint64_t>(
          span<double>(t.bicubic_derivative_energy, t.bicubic_coefficient_count)),
      NativeArrayView<double, std::uint64_t>(
          span<double>(t.bicubic_derivative_loss, t.bicubic_coefficient_count)),
      NativeArrayView<double, std::uint64_t>(
          span<double>(t.bicubic_mixed_derivative, t.bicubic_coefficient_count)),
      NativeArrayView<double, std::uint64_t>(span<double>(
          t.bicubic_polynomial_coefficients, t.bicubic_polynomial_coefficient_count)),
      NativeArrayView<double, std::uint64_t>(span<double>(t.cubic_values, t.cubic_coefficient_count)),
      NativeArrayView<double, std::uint64_t>(
          span<double>(t.cubic_node_derivatives, t.cubic_coefficient_count)),
      constants);
}

c8mi_native_result convert(ck::NativeQueryResult const& r) {
  c8mi_native_result out{};
  out.status = static_cast<std::uint32_t>(r.status_);
  out.newton_iterations = r.newtonIterations_;
  out.bisection_iterations = r.bisectionIterations_;
  out.value = r.value_;
  return out;
}

}  // namespace

static_assert(sizeof(c8mi_moliere_snapshot) == sizeof(ck::MoliereSnapshot));
static_assert(offsetof(c8mi_moliere_snapshot, components) == offsetof(ck::MoliereSnapshot, components_));
static_assert(offsetof(c8mi_moliere_snapshot, c1) == offsetof(ck::MoliereSnapshot, c1_));
static_assert(offsetof(c8mi_moliere_snapshot, c2_large) == offsetof(ck::MoliereSnapshot, c2Large_));
static_assert(offsetof(c8mi_moliere_snapshot, c1_integral_large) ==
              offsetof(ck::MoliereSnapshot, c1IntegralLarge_));

static ck::MoliereSnapshot as_ref(c8mi_moliere_snapshot const* s) {
  ck::MoliereSnapshot r;
  std::memcpy(&r, s, sizeof r);
  return r;
}

static_assert(sizeof(c8mi_moliere_interpolation) == sizeof(ck::MoliereInterpolationTable));
static_assert(offsetof(c8mi_moliere_interpolation, initial_guess_delta) ==
              offsetof(ck::MoliereInterpolationTable, initialGuessDelta_));
static_assert(sizeof(c8mi_moliere_cubic) == sizeof(ck::MoliereCubicPolynomial));

// The reference's view (make_moliere_interpolation_view) over a copy of a c8mi
// table in the reference's own record, kept while the same table is probed.
static ck::MoliereInterpolationView as_view(c8mi_moliere_interpolation const* t) {
  static std::vector<unsigned char> storage(sizeof(ck::MoliereInterpolationTable));
  static bool filled = false;
  auto* table = reinterpret_cast<ck::MoliereInterpolationTable*>(storage.data());
  if (!filled || std::memcmp(storage.data(), t, sizeof *t) != 0) {
    std::memcpy(storage.data(), t, sizeof *t);
    filled = true;
  }
  return ck::make_moliere_interpolation_view(
      ck::NativeArrayView<ck::MoliereCubicPolynomial>{table->polynomials_},
      ck::NativeArrayView<double>{table->initialGuessDelta_});
}

extern "C" {

// f1, f2 and their integrals (analytical functions, no interpolation view).
void ref_moliere_functions(c8mi_moliere_snapshot const* s, double x, double out[4]) {
  auto const r = as_ref(s);
  out[0] = ck::moliere_detail::f1_m(r, x);
  out[1] = ck::moliere_detail::f2_m(r, x);
  out[2] = ck::moliere_detail::integral_f1_m(r, x);
  out[3] = ck::moliere_detail::integral_f2_m(r, x);
}

double ref_moliere_solve_b(double rhs) { return ck::moliere_detail::solve_moliere_b(rhs); }

double ref_moliere_normal_ppf(double p) { return ck::moliere_detail::normal_ppf(p); }

int ref_moliere_angle(c8mi_moliere_snapshot const* s, double grammage, double e_mev, double u1, double u2,
                      uint32_t* iterations, double* angle) {
  auto const r = as_ref(s);
  auto const a = ck::sample_moliere_scattering_angle_2d(r, grammage, e_mev, 0.9 * e_mev, u1, u2, true);
  *iterations = a.iterations_;
  *angle = a.angleRad_;
  return static_cast<int>(a.status_);
}

// The same with the interpolation view of a table.
void ref_moliere_interpolated_functions(c8mi_moliere_snapshot const* s, c8mi_moliere_interpolation const* t,
                                        double x, double out[4]) {
  auto const r = as_ref(s);
  auto const v = as_view(t);
  out[0] = ck::moliere_detail::f1_m(r, v, x);
  out[1] = ck::moliere_detail::f2_m(r, v, x);
  out[2] = ck::moliere_detail::integral_f1_m(r, v, x);
  out[3] = ck::moliere_detail::integral_f2_m(r, v, x);
}

int ref_moliere_angle_interpolated(c8mi_moliere_snapshot const* s, c8mi_moliere_interpolation const* t,
                                   double grammage, double e_mev, double u1, double u2, uint32_t* iterations,
                                   double* angle) {
  auto const r = as_ref(s);
  auto const a = ck::sample_moliere_scattering_angle_2d(r, as_view(t), grammage, e_mev, 0.9 * e_mev, u1, u2, true);
  *iterations = a.iterations_;
  *angle = a.angleRad_;
  return static_cast<int>(a.status_);
}

int ref_moliere_direction(double const start[3], double const tracked[3], double angle, double azimuth_uniform,
                          double out[3]) {
  double s3[3] = {start[0], start[1], start[2]}, t3[3] = {tracked[0], tracked[1], tracked[2]};
  auto const d = ck::apply_moliere_direction_cpu_step_compatible(s3, t3, angle, azimuth_uniform);
  for (int k = 0; k < 3; ++k) out[k] = d.direction_[k];
  return static_cast<int>(d.status_);
}


void ref_philox4x32_10(std::uint32_t const counter[4], std::uint32_t const key[2],
                       std::uint32_t out[4]) {
  ck::PhiloxCounter c{{counter[0], counter[1], counter[2], counter[3]}};
  ck::PhiloxKey k{{key[0], key[1]}};
  auto const r = ck::philox4x32_10(c, k);
  for (int i = 0; i < 4; ++i) out[i] = r.words_[i];
}

void ref_random_words(std::uint64_t seed, std::uint64_t shower, std::uint64_t history,
                      std::uint64_t step, std::uint32_t process, std::uint64_t draw,
                      std::uint32_t out[4]) {
  ck::RandomNumberKey key{seed, shower, history, step, process, draw};
  auto const r = ck::random_words(key);
  for (int i = 0; i < 4; ++i) out[i] = r.words_[i];
}

double ref_uniform_open01(std::uint64_t seed, std::uint64_t shower, std::uint64_t history,
                          std::uint64_t step, std::uint32_t process, std::uint64_t draw,
                          std::uint32_t lane) {
  ck::RandomNumberKey key{seed, shower, history, step, process, draw};
  return ck::uniform_open01(key, lane);
}

void ref_native_queries(c8mi_native_table_set const* table, c8mi_native_query const* queries,
                        std::uint64_t count, c8mi_native_result* out) {
  auto const view = make_view(*table);
  for (std::uint64_t i = 0; i < count; ++i) {
    auto const& q = queries[i];
    ck::NativeQueryResult r{};
    switch (q.kind) {
      case C8MI_NATIVE_TOTAL_RATE:
        r = ck::query_proposal_native_total_rate(view, q.pdg_id, q.energy_mev);
        break;
      case C8MI_NATIVE_SELECTION_TOTAL_RATE:
        r = ck::query_proposal_native_selection_total_rate(view, q.pdg_id, q.energy_mev);
        break;
      case C8MI_NATIVE_RATE:
        r = ck::query_proposal_native_rate(view, q.pdg_id, q.process_id, q.component_hash,
                                           q.energy_mev);
        break;
      case C8MI_NATIVE_DEDX:
        r = ck::query_proposal_native_dedx(view, q.pdg_id, q.energy_mev);
        break;
      case C8MI_NATIVE_LOSS_FRACTION:
      case C8MI_NATIVE_PHOTON_SPLIT:
        r = ck::query_proposal_native_loss_fraction(view, q.pdg_id, q.process_id,
                                                    q.component_hash, q.energy_mev,
                                                    q.argument,
                                                    q.kind == C8MI_NATIVE_PHOTON_SPLIT);
        break;
      case C8MI_NATIVE_RANGE:
        r = ck::query_proposal_native_range(view, q.pdg_id, q.energy_mev);
        break;
      case C8MI_NATIVE_ENERGY_AFTER_LOSS:
        r = ck::query_proposal_native_energy_after_continuous_loss(view, q.pdg_id,
                                                                   q.energy_mev, q.argument);
        break;
      case C8MI_NATIVE_ENERGY:
        r = ck::query_proposal_native_energy(view, q.pdg_id, q.argument);
        break;
      case C8MI_NATIVE_CUMULATIVE_RATE: {
        auto const* column = ck::proposal_native_detail::find_dndx(view, q.pdg_id, q.process_id,
                                                                   q.component_hash);
        if (!column)
          r = {ck::NativeQueryStatus::ColumnNotFound, 0, 0, 0.};
        else
          r = ck::query_proposal_native_cumulative_rate(view, *column, q.energy_mev,
                                                        q.argument);
        break;
      }
      default:
        r = {ck::NativeQueryStatus::InvalidView, 0, 0, 0.};
    }
    out[i] = convert(r);
  }
}


// ---- environment: the c8mi tree converted to the reference's data view ----

namespace {

struct RefEnvironment {
  std::vector<ck::EnvironmentNode> nodes;
  std::vector<ck::EnvironmentShape> shapes;
  std::vector<std::int32_t> children;
  std::vector<ck::EnvironmentModel> models;
  ck::EnvironmentBorrowedView view;
  ck::EnvironmentQueryConfig config;

  explicit RefEnvironment(c8mi_environment const& e) {
    for (std::uint32_t i = 0; i < e.node_count; ++i) {
      auto const& n = e.nodes[i];
      ck::EnvironmentNode rn{};
      rn.parentIndex_ = n.parent;
      rn.modelIndex_ = n.model;
      rn.childOffset_ = n.child_begin;
      rn.childCount_ = n.child_count;
      nodes.push_back(rn);
      ck::EnvironmentShape sh{};
      sh.kind_ = n.shape_kind == C8MI_SHAPE_UNIVERSE ? ck::EnvironmentShape::Kind::Universe
                                                     : ck::EnvironmentShape::Kind::Sphere;
      for (int a = 0; a < 3; ++a) sh.centerM_[a] = n.center_m[a];
      sh.radiusM_ = n.radius_m;
      shapes.push_back(sh);
    }
    children.assign(e.children, e.children + e.child_index_count);
    for (std::uint32_t i = 0; i < e.model_count; ++i) {
      auto const& m = e.models[i];
      ck::EnvironmentModel rm{};
      rm.densityKind_ = m.density_kind == C8MI_DENSITY_HOMOGENEOUS ? ck::EnvironmentModel::Density::Homogeneous
                        : m.density_kind == C8MI_DENSITY_SLIDING_PLANAR_EXPONENTIAL
                            ? ck::EnvironmentModel::Density::SlidingPlanarExponential
                            : ck::EnvironmentModel::Density::FlatExponential;
      rm.referenceDensityGPerCm3_ = m.reference_density_g_per_cm3;
      rm.scaleLengthM_ = m.scale_length_m;
      rm.inverseScaleLengthPerM_ = m.inverse_scale_length_per_m;
      rm.referenceHeightM_ = m.reference_height_m;
      for (int a = 0; a < 3; ++a) {
        rm.anchorM_[a] = m.anchor_m[a];
        rm.axis_[a] = m.axis[a];
        rm.magneticFieldT_[a] = m.magnetic_field_t[a];
      }
      rm.mediumId_ = m.medium_id;
      models.push_back(rm);
    }
    ck::EnvironmentDataView d{};
    d.nodes_ = nodes.data();
    d.shapes_ = shapes.data();
    d.children_ = children.data();
    d.models_ = models.data();
    d.nodeCount_ = static_cast<std::uint32_t>(nodes.size());
    d.childCount_ = static_cast<std::uint32_t>(children.size());
    d.modelCount_ = static_cast<std::uint32_t>(models.size());
    view = ck::EnvironmentBorrowedView(d);
    config.maximumMagneticDeflectionRad_ = e.maximum_magnetic_deflection_rad;
  }
};

}  // namespace

int ref_em_environment_query(c8mi_environment const* e, double const position[3],
                             double const direction[3], std::int32_t node, std::int32_t* out_node,
                             double* out_density) {
  RefEnvironment env(*e);
  double const p[3] = {position[0], position[1], position[2]};
  auto const r = ck::query_environment(env.view, env.config, p, direction, node);
  *out_node = r.nodeIndex_;
  *out_density = r.densityGPerCm3_;
  return static_cast<int>(r.status_);
}

int ref_em_boundary(c8mi_environment const* e, double const position[3], double const direction[3],
                    std::int32_t node, double* distance, std::int32_t* next_node) {
  RefEnvironment env(*e);
  double const p[3] = {position[0], position[1], position[2]};
  double const d[3] = {direction[0], direction[1], direction[2]};
  auto const r = ck::distance_to_environment_boundary(env.view, env.config, p, d, node);
  *distance = r.distanceM_;
  *next_node = r.nextNode_;
  return static_cast<int>(r.status_);
}

int ref_em_grammage(c8mi_environment const* e, std::int32_t node, double const position[3],
                    double const direction[3], double distance, double* out) {
  RefEnvironment env(*e);
  double const p[3] = {position[0], position[1], position[2]};
  double const d[3] = {direction[0], direction[1], direction[2]};
  auto const r = ck::environment_grammage(env.view, env.config, node, p, d, distance);
  *out = r.value_;
  return static_cast<int>(r.status_);
}

int ref_em_arclength(c8mi_environment const* e, std::int32_t node, double const position[3],
                     double const direction[3], double grammage, double* out) {
  RefEnvironment env(*e);
  double const p[3] = {position[0], position[1], position[2]};
  double const d[3] = {direction[0], direction[1], direction[2]};
  auto const r = ck::environment_arclength(env.view, env.config, node, p, d, grammage);
  *out = r.value_;
  return static_cast<int>(r.status_);
}

int ref_em_pair_lpm(c8mi_photon_pair_lpm const* s, std::uint64_t comp, double e_mev, double x,
                    double rho, double* out) {
  ck::PhotonPairLpmSnapshot snap{};
  std::memcpy(&snap, s, sizeof snap);
  auto const r = ck::photon_pair_lpm_suppression_factor(snap, comp, e_mev, x, rho);
  *out = r.survivalProbability_;
  return static_cast<int>(r.status_);
}

int ref_em_brems_lpm(c8mi_brems_lpm const* s, std::uint64_t comp, double e_mev, double v, double rho,
                     double* out) {
  ck::BremsLpmSnapshot snap{};
  std::memcpy(&snap, s, sizeof snap);
  auto const r = ck::brems_lpm_suppression_factor(snap, comp, e_mev, v, rho);
  *out = r.survivalProbability_;
  return static_cast<int>(r.status_);
}

int ref_em_thinning(c8mi_em_thinning const* c, double pe, double pw, double e1, double e2, double u1,
                    double u2, std::uint32_t* keep, double* w1, double* w2) {
  ck::EmThinningConfig cfg{};
  cfg.thresholdGeV_ = c->threshold_gev;
  cfg.maximumWeight_ = c->maximum_weight;
  cfg.enabled_ = c->enabled;
  cfg.eraseZeroWeight_ = c->erase_zero_weight;
  auto const r = ck::apply_em_thinning(cfg, pe, pw, e1, e2, u1, u2);
  *keep = r.keepMask_;
  *w1 = r.firstWeight_;
  *w2 = r.secondWeight_;
  return static_cast<int>(r.status_);
}

int ref_em_pair_split(double e_gev, double x, double* normalized, double* decoded) {
  if (!ck::encode_photon_pair_normalized_split(e_gev, x, *normalized)) return 0;
  return ck::decode_photon_pair_normalized_split(e_gev, *normalized, *decoded) ? 1 : 0;
}


/* The reference's content hash of a table set with its metadata
 * (calculate_proposal_native_hash, ProposalNativeTable.cpp:570-666) and its
 * validation (validate_proposal_native_table, :234-568; returns 0 when valid,
 * else 1 with the message in `message`). */
static ck::ProposalNativeTableSet owning_set(c8mi_native_table_set const* t, c8mi_native_table_meta const* m) {
  ck::ProposalNativeTableSet s;
  s.formatVersion_ = m->format_version;
  s.proposalVersion_ = m->proposal_version;
  s.cubicInterpolationVersion_ = m->cubic_interpolation_version;
  std::memcpy(&s.constants_, &t->constants, sizeof(s.constants_));
  s.stochasticCutMeV_ = m->stochastic_cut_mev;
  for (uint64_t i = 0; i < m->interaction_identity_count; ++i) {
    ck::NativeInteractionIdentity id{};
    id.pdgId_ = m->interaction_identities[i].pdg_id;
    id.reserved_ = m->interaction_identities[i].reserved;
    id.mediumHash_ = m->interaction_identities[i].medium_hash;
    id.interactionHash_ = m->interaction_identities[i].interaction_hash;
    s.interactionIdentities_.push_back(id);
  }
  auto copy = [](auto& out, auto const* in, std::size_t n) {
    out.resize(n);
    if (n) std::memcpy(out.data(), in, n * sizeof(out[0]));
  };
  copy(s.dndxColumns_, t->dndx_columns, t->dndx_column_count);
  copy(s.totalRateColumns_, t->total_rate_columns, t->total_rate_column_count);
  copy(s.selectionColumnIndices_, t->selection_column_indices, t->dndx_column_count);
  copy(s.dedxColumns_, t->dedx_columns, t->dedx_column_count);
  copy(s.dedxAccumulationIndices_, t->dedx_accumulation_indices, t->dedx_column_count);
  copy(s.utilityColumns_, t->utility_columns, t->utility_column_count);
  copy(s.bicubicValues_, t->bicubic_values, t->bicubic_coefficient_count);
  copy(s.bicubicDerivativeEnergy_, t->bicubic_derivative_energy, t->bicubic_coefficient_count);
  copy(s.bicubicDerivativeLoss_, t->bicubic_derivative_loss, t->bicubic_coefficient_count);
  copy(s.bicubicMixedDerivative_, t->bicubic_mixed_derivative, t->bicubic_coefficient_count);
  copy(s.bicubicPolynomialCoefficients_, t->bicubic_polynomial_coefficients, t->bicubic_polynomial_coefficient_count);
  copy(s.cubicValues_, t->cubic_values, t->cubic_coefficient_count);
  copy(s.cubicNodeDerivatives_, t->cubic_node_derivatives, t->cubic_coefficient_count);
  return s;
}

int ref_native_table_hash(c8mi_native_table_set const* t, c8mi_native_table_meta const* m, uint8_t out[32]) {
  auto const digest = ck::calculate_proposal_native_hash(owning_set(t, m));
  std::memcpy(out, digest.data(), 32);
  return 0;
}

int ref_native_table_validate(c8mi_native_table_set const* t, c8mi_native_table_meta const* m, char* message,
                              uint64_t capacity) {
  try {
    ck::validate_proposal_native_table(owning_set(t, m));
    return 0;
  } catch (std::exception const& e) {
    if (message && capacity) {
      std::strncpy(message, e.what(), capacity - 1);
      message[capacity - 1] = '\0';
    }
    return 1;
  }
}

int ref_sha256(uint8_t const* data, uint64_t size, uint8_t out[32]) {
  std::vector<std::uint8_t> bytes(data, data + size);
  auto const digest = ck::sha256(bytes);
  std::memcpy(out, digest.data(), 32);
  return 0;
}

}  // extern "C"
