// TEST INFRASTRUCTURE ONLY: a C-ABI probe over the reference's own header-only
// functions, compiled in place from /root/reference by oracle/ref.mk into
// oracle/_ref/libref.so. It lets the tests check this repository's C
// restatement (and the HIP kernels) against the reference code itself on the
// same inputs. Only headers that compile unmodified are included; nothing of
// the reference is copied here.
#include <corsika/kokkos/em/random/Philox.hpp>
#include <corsika/kokkos/em/tables/ProposalNativeQueries.hpp>

#include <cstddef>
#include <cstdint>

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

extern "C" {

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

}  // extern "C"
