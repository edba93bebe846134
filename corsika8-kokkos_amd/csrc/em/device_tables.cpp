#include "em/device_tables.hpp"

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace c8mi::em {

namespace {

void require(bool condition, char const* message) {
  if (!condition) throw std::invalid_argument(std::string("native tables: ") + message);
}

void check_axis(c8mi_native_axis const& a, std::uint32_t expected_nodes, char const* what) {
  require(a.type <= C8MI_AXIS_EXPONENTIAL_MINUS_ONE, what);
  require(a.nodes >= 2 && (expected_nodes == 0 || a.nodes == expected_nodes), what);
  require(std::isfinite(a.low) && std::isfinite(a.high) && std::isfinite(a.step) &&
              a.high > a.low && a.step > 0.,
          what);
  require(a.type == C8MI_AXIS_LINEAR || a.low > 0., what);
}

void check_cubic(c8mi_native_cubic const& s, std::uint64_t coefficients, char const* what) {
  require(s.coefficient_count >= 2 && s.coefficient_offset <= coefficients &&
              s.coefficient_count <= coefficients - s.coefficient_offset,
          what);
  check_axis(s.axis, s.coefficient_count, what);
  require(s.has_value_transform <= 1u, what);
  if (s.has_value_transform) {
    auto const& t = s.value_transform;
    require(t.type <= C8MI_AXIS_EXPONENTIAL_MINUS_ONE && std::isfinite(t.low) &&
                std::isfinite(t.step) && t.step > 0. &&
                (t.type == C8MI_AXIS_LINEAR || t.low > 0.),
            what);
  }
}

}  // namespace

void validate_table_bounds(c8mi_native_table_set const& t) {
  require(t.dndx_column_count > 0 && t.dndx_columns != nullptr, "no dN/dX columns");
  require(t.total_rate_column_count > 0 && t.total_rate_columns != nullptr,
          "no total-rate columns");
  require(t.selection_column_indices != nullptr, "no selection index");
  require(t.dedx_column_count == 0 ||
              (t.dedx_columns != nullptr && t.dedx_accumulation_indices != nullptr),
          "dE/dX columns without accumulation index");
  require(t.utility_column_count == 0 || t.utility_columns != nullptr, "utility columns");
  require(t.bicubic_coefficient_count > 0 && t.bicubic_values && t.bicubic_derivative_energy &&
              t.bicubic_derivative_loss && t.bicubic_mixed_derivative,
          "bicubic arrays");
  require(t.bicubic_polynomial_coefficient_count == 0 || t.bicubic_polynomial_coefficients,
          "polynomial coefficients");
  require(t.cubic_coefficient_count == 0 || (t.cubic_values && t.cubic_node_derivatives),
          "cubic arrays");

  for (std::uint32_t i = 0; i < t.dndx_column_count; ++i) {
    auto const& c = t.dndx_columns[i];
    require(t.selection_column_indices[i] < t.dndx_column_count, "selection index out of range");
    if (i > 0) {
      auto const& p = t.dndx_columns[i - 1];
      bool const ordered =
          p.pdg_id < c.pdg_id ||
          (p.pdg_id == c.pdg_id &&
           (p.process_id < c.process_id ||
            (p.process_id == c.process_id && p.component_hash < c.component_hash)));
      require(ordered, "dN/dX columns are not in canonical (pid, process, component) order");
    }
    if (c.rate_model != C8MI_RATE_BICUBIC) continue;
    auto const& s = c.spline;
    require(s.rows >= 2 && s.columns >= 2, "bicubic dimensions");
    check_axis(s.energy_axis, s.rows, "dN/dX energy axis");
    check_axis(s.loss_axis, s.columns, "dN/dX loss axis");
    std::uint64_t const count = static_cast<std::uint64_t>(s.rows) * s.columns;
    require(s.coefficient_offset <= t.bicubic_coefficient_count &&
                count <= t.bicubic_coefficient_count - s.coefficient_offset,
            "bicubic descriptor out of bounds");
    if (s.polynomial_coefficient_offset != C8MI_NO_POLYNOMIAL_COEFFICIENTS) {
      std::uint64_t const cells = static_cast<std::uint64_t>(s.rows - 1u) * (s.columns - 1u);
      std::uint64_t const available = t.bicubic_polynomial_coefficient_count;
      require(s.polynomial_coefficient_offset <= available &&
                  cells <= (available - s.polynomial_coefficient_offset) / 16u,
              "polynomial descriptor out of bounds");
    }
  }
  for (std::uint32_t i = 0; i < t.total_rate_column_count; ++i) {
    check_cubic(t.total_rate_columns[i].spline, t.cubic_coefficient_count, "total-rate spline");
    require(i == 0 || t.total_rate_columns[i - 1].pdg_id < t.total_rate_columns[i].pdg_id,
            "total-rate columns are not sorted by PID");
  }
  for (std::uint32_t i = 0; i < t.dedx_column_count; ++i) {
    require(t.dedx_accumulation_indices[i] < t.dedx_column_count, "dE/dX index out of range");
    check_cubic(t.dedx_columns[i].spline, t.cubic_coefficient_count, "dE/dX spline");
  }
  for (std::uint32_t i = 0; i < t.utility_column_count; ++i)
    check_cubic(t.utility_columns[i].spline, t.cubic_coefficient_count, "utility spline");
}

DeviceTables::DeviceTables(c8mi_native_table_set const& host) {
  validate_table_bounds(host);

  struct Part {
    void const* source;
    std::size_t bytes;
    std::size_t offset;
  };
  std::vector<Part> parts;
  std::size_t total = 0;
  auto add = [&](void const* source, std::size_t bytes) {
    total = (total + 255u) & ~std::size_t{255u};
    parts.push_back({source, bytes, total});
    total += bytes;
    return parts.size() - 1;
  };
  std::uint32_t const nd = host.dndx_column_count;
  std::size_t const i_dndx = add(host.dndx_columns, nd * sizeof(c8mi_native_dndx_column));
  std::size_t const i_total = add(host.total_rate_columns,
                                  host.total_rate_column_count * sizeof(c8mi_native_total_rate_column));
  std::size_t const i_sel = add(host.selection_column_indices, nd * sizeof(std::uint32_t));
  std::size_t const i_dedx =
      add(host.dedx_columns, host.dedx_column_count * sizeof(c8mi_native_dedx_column));
  std::size_t const i_dedx_index =
      add(host.dedx_accumulation_indices, host.dedx_column_count * sizeof(std::uint32_t));
  std::size_t const i_util =
      add(host.utility_columns, host.utility_column_count * sizeof(c8mi_native_utility_column));
  std::size_t const nb = host.bicubic_coefficient_count * sizeof(double);
  std::size_t const i_b0 = add(host.bicubic_values, nb);
  std::size_t const i_b1 = add(host.bicubic_derivative_energy, nb);
  std::size_t const i_b2 = add(host.bicubic_derivative_loss, nb);
  std::size_t const i_b3 = add(host.bicubic_mixed_derivative, nb);
  std::size_t const i_poly =
      add(host.bicubic_polynomial_coefficients, host.bicubic_polynomial_coefficient_count * sizeof(double));
  std::size_t const nc = host.cubic_coefficient_count * sizeof(double);
  std::size_t const i_c0 = add(host.cubic_values, nc);
  std::size_t const i_c1 = add(host.cubic_node_derivatives, nc);

  std::vector<unsigned char> staging(total, 0);
  for (auto const& p : parts)
    if (p.bytes) std::memcpy(staging.data() + p.offset, p.source, p.bytes);
  storage_.allocate(total);
  C8MI_HIP(hipMemcpy(storage_.data(), staging.data(), total, hipMemcpyHostToDevice));

  unsigned char* base = storage_.data();
  auto at = [&](std::size_t i) -> void* {
    return parts[i].bytes ? static_cast<void*>(base + parts[i].offset) : nullptr;
  };
  view_ = host;
  view_.dndx_columns = static_cast<c8mi_native_dndx_column const*>(at(i_dndx));
  view_.total_rate_columns = static_cast<c8mi_native_total_rate_column const*>(at(i_total));
  view_.selection_column_indices = static_cast<std::uint32_t const*>(at(i_sel));
  view_.dedx_columns = static_cast<c8mi_native_dedx_column const*>(at(i_dedx));
  view_.dedx_accumulation_indices = static_cast<std::uint32_t const*>(at(i_dedx_index));
  view_.utility_columns = static_cast<c8mi_native_utility_column const*>(at(i_util));
  view_.bicubic_values = static_cast<double const*>(at(i_b0));
  view_.bicubic_derivative_energy = static_cast<double const*>(at(i_b1));
  view_.bicubic_derivative_loss = static_cast<double const*>(at(i_b2));
  view_.bicubic_mixed_derivative = static_cast<double const*>(at(i_b3));
  view_.bicubic_polynomial_coefficients = static_cast<double const*>(at(i_poly));
  view_.cubic_values = static_cast<double const*>(at(i_c0));
  view_.cubic_node_derivatives = static_cast<double const*>(at(i_c1));
}

}  // namespace c8mi::em
