#include "em/device_tables.hpp"

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
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

namespace {

// The arrays of a table set packed into one 256-byte-aligned image; `at`
// gives each array's offset (npos: empty).
struct Packed {
  static constexpr std::size_t npos = ~std::size_t{0};
  std::vector<unsigned char> image;
  std::size_t at[13];
};

Packed pack(c8mi_native_table_set const& host) {
  Packed p;
  std::size_t total = 0;
  struct Part {
    void const* source;
    std::size_t bytes;
  };
  std::uint32_t const nd = host.dndx_column_count;
  std::size_t const nb = host.bicubic_coefficient_count * sizeof(double);
  std::size_t const nc = host.cubic_coefficient_count * sizeof(double);
  Part const parts[13] = {
      {host.dndx_columns, nd * sizeof(c8mi_native_dndx_column)},
      {host.total_rate_columns, host.total_rate_column_count * sizeof(c8mi_native_total_rate_column)},
      {host.selection_column_indices, nd * sizeof(std::uint32_t)},
      {host.dedx_columns, host.dedx_column_count * sizeof(c8mi_native_dedx_column)},
      {host.dedx_accumulation_indices, host.dedx_column_count * sizeof(std::uint32_t)},
      {host.utility_columns, host.utility_column_count * sizeof(c8mi_native_utility_column)},
      {host.bicubic_values, nb},
      {host.bicubic_derivative_energy, nb},
      {host.bicubic_derivative_loss, nb},
      {host.bicubic_mixed_derivative, nb},
      {host.bicubic_polynomial_coefficients, host.bicubic_polynomial_coefficient_count * sizeof(double)},
      {host.cubic_values, nc},
      {host.cubic_node_derivatives, nc}};
  for (int i = 0; i < 13; ++i) {
    total = (total + 255u) & ~std::size_t{255u};
    p.at[i] = parts[i].bytes ? total : Packed::npos;
    total += parts[i].bytes;
  }
  p.image.assign(total, 0);
  for (int i = 0; i < 13; ++i)
    if (parts[i].bytes) std::memcpy(p.image.data() + p.at[i], parts[i].source, parts[i].bytes);
  return p;
}

// The table set's view with every array at its offset from `base`.
c8mi_native_table_set rebase(c8mi_native_table_set const& host, Packed const& p, unsigned char* base) {
  auto at = [&](int i) -> void* { return p.at[i] == Packed::npos ? nullptr : static_cast<void*>(base + p.at[i]); };
  c8mi_native_table_set v = host;
  v.dndx_columns = static_cast<c8mi_native_dndx_column const*>(at(0));
  v.total_rate_columns = static_cast<c8mi_native_total_rate_column const*>(at(1));
  v.selection_column_indices = static_cast<std::uint32_t const*>(at(2));
  v.dedx_columns = static_cast<c8mi_native_dedx_column const*>(at(3));
  v.dedx_accumulation_indices = static_cast<std::uint32_t const*>(at(4));
  v.utility_columns = static_cast<c8mi_native_utility_column const*>(at(5));
  v.bicubic_values = static_cast<double const*>(at(6));
  v.bicubic_derivative_energy = static_cast<double const*>(at(7));
  v.bicubic_derivative_loss = static_cast<double const*>(at(8));
  v.bicubic_mixed_derivative = static_cast<double const*>(at(9));
  v.bicubic_polynomial_coefficients = static_cast<double const*>(at(10));
  v.cubic_values = static_cast<double const*>(at(11));
  v.cubic_node_derivatives = static_cast<double const*>(at(12));
  return v;
}

}  // namespace

DeviceTables::DeviceTables(c8mi_native_table_set const& host) {
  validate_table_bounds(host);
  Packed const p = pack(host);
  storage_.allocate(p.image.size());
  C8MI_HIP(hipMemcpy(storage_.data(), p.image.data(), p.image.size(), hipMemcpyHostToDevice));
  view_ = rebase(host, p, storage_.data());
}

HostTables::HostTables(c8mi_native_table_set const& host) {
  validate_table_bounds(host);
  Packed p = pack(host);
  storage_ = std::move(p.image);
  view_ = rebase(host, p, storage_.data());
}

}  // namespace c8mi::em
