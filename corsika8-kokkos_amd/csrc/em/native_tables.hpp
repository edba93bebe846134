// Device evaluation of format-v7 native tables: the PROPOSAL-native query API
// of corsika/detail/kokkos/em/tables/ProposalNativeQueries.inl:88-1428
// (cubic / bicubic Hermite splines on linear, exponential and expm1 axes, the
// six-point finite-difference loss derivative, Boost-1.79-compatible Newton
// with PROPOSAL's bisection fallbacks, kinematic limits and the direct rate
// models). The table set is c8mi_native_table_set with device pointers.
//
// Status codes, iteration counts and branch order follow the reference; values
// agree with the FP64 host evaluation to the last few ulps of the device
// log/exp/pow (the test bar is 1e-9 relative).
#pragma once

#include <hip/hip_runtime.h>

#include "em/hd.hpp"

#include <cmath>
#include <cstdint>

#include "c8mi/c8mi_tables.h"

namespace c8mi::em {

using Tables = c8mi_native_table_set;

struct QueryResult {
  std::uint32_t status;
  std::uint32_t newton;
  std::uint32_t bisection;
  double value;
};

C8MI_HD QueryResult qr(std::uint32_t status, double value = 0., std::uint32_t newton = 0,
              std::uint32_t bisection = 0) {
  return {status, newton, bisection, value};
}

namespace detail {

constexpr double kLn2 = 0.693147180559945309417232121458176568;

C8MI_HD double axis_transform(c8mi_native_axis const& a, double x) {
  switch (a.type) {
    case C8MI_AXIS_LINEAR: return (x - a.low) / a.step;
    case C8MI_AXIS_EXPONENTIAL: return log(x / a.low) / a.step;
    case C8MI_AXIS_EXPONENTIAL_MINUS_ONE: return (log1p(x / a.low) - kLn2) / a.step;
  }
  return __builtin_nan("");
}

C8MI_HD double axis_back_transform(c8mi_native_axis const& a, double c) {
  switch (a.type) {
    case C8MI_AXIS_LINEAR: return c * a.step + a.low;
    case C8MI_AXIS_EXPONENTIAL: return a.low * exp(c * a.step);
    case C8MI_AXIS_EXPONENTIAL_MINUS_ONE: return a.low * expm1(c * a.step + kLn2);
  }
  return __builtin_nan("");
}

C8MI_HD double axis_derivative(c8mi_native_axis const& a, double x) {
  switch (a.type) {
    case C8MI_AXIS_LINEAR: return 1. / a.step;
    case C8MI_AXIS_EXPONENTIAL: return 1. / (x * a.step);
    case C8MI_AXIS_EXPONENTIAL_MINUS_ONE: return 1. / (a.step * (x + a.low));
  }
  return __builtin_nan("");
}

C8MI_HD double axis_back_derivative(c8mi_native_axis const& a, double c) {
  switch (a.type) {
    case C8MI_AXIS_LINEAR: return a.step;
    case C8MI_AXIS_EXPONENTIAL: return a.low * a.step * exp(c * a.step);
    case C8MI_AXIS_EXPONENTIAL_MINUS_ONE: return a.low * a.step * exp(c * a.step + kLn2);
  }
  return __builtin_nan("");
}

// Interval below `coordinate`, clamped to [0, nodes - 2]. A NaN coordinate
// maps to 0 (the host form indexes with (uint32)-1 there); the fraction is NaN
// either way, so the interpolated value is NaN as on the host, and no load
// leaves the spline's coefficients.
C8MI_HD std::uint32_t interval_node(double coordinate, std::uint32_t nodes) {
  if (!(coordinate >= 0.)) return 0;
  if (coordinate >= static_cast<double>(nodes - 2u)) return nodes - 2u;
  return static_cast<std::uint32_t>(coordinate);
}

C8MI_HD void hermite(double x, double b[4]) {
  double const x2 = x * x, x3 = x2 * x;
  b[0] = 1. - 3. * x2 + 2. * x3;
  b[1] = 3. * x2 - 2. * x3;
  b[2] = x - 2. * x2 + x3;
  b[3] = -x2 + x3;
}

C8MI_HD void hermite_derivative(double x, double b[4]) {
  double const x2 = x * x;
  b[0] = -6. * x + 6. * x2;
  b[1] = 6. * x - 6. * x2;
  b[2] = 1. - 4. * x + 3. * x2;
  b[3] = -2. * x + 3. * x2;
}

// Energy-axis part of a bicubic evaluation. Within one query the energy is
// fixed while the loss varies (root finding, finite differences), so the
// energy coordinate, interval and fraction are computed once per query; the
// values are the ones every evaluation would compute.
struct EnergyRow {
  double x0;
  std::uint32_t n0;
  double f0;
};

C8MI_HD EnergyRow energy_row(c8mi_native_bicubic const& s, double e) {
  double const x0 = axis_transform(s.energy_axis, e);
  std::uint32_t const n0 = interval_node(x0, s.rows);
  return {x0, n0, x0 - n0};
}

// Four coefficient planes, column-major per spline: (row, column) at
// offset + row + rows * column.
C8MI_HD double bicubic_sum(Tables const& t, c8mi_native_bicubic const& s, EnergyRow const& er,
                           double loss, bool derivative) {
  double const x1 = axis_transform(s.loss_axis, loss);
  std::uint32_t const n0 = er.n0;
  std::uint32_t const n1 = interval_node(x1, s.columns);
  double h0[4], h1[4];
  hermite(er.f0, h0);
  if (derivative) hermite_derivative(x1 - n1, h1);
  else hermite(x1 - n1, h1);
  double r = 0.;
#pragma unroll
  for (std::uint32_t a = 0; a < 4; ++a) {
#pragma unroll
    for (std::uint32_t b = 0; b < 4; ++b) {
      double const* plane = (a >= 2 && b < 2)   ? t.bicubic_derivative_energy
                            : (a < 2 && b >= 2) ? t.bicubic_derivative_loss
                            : (a >= 2 && b >= 2) ? t.bicubic_mixed_derivative
                                                 : t.bicubic_values;
      std::uint64_t const local =
          static_cast<std::uint64_t>(n0 + (a & 1u)) +
          static_cast<std::uint64_t>(s.rows) * static_cast<std::uint64_t>(n1 + (b & 1u));
      r += h0[a] * plane[s.coefficient_offset + local] * h1[b];
    }
  }
  return derivative ? r / s.loss_axis.step : r;
}

C8MI_HD double dot4(double a0, double a1, double a2, double a3, double b0, double b1,
           double b2, double b3) {
  double const p0 = a0 * b0, p1 = a1 * b1, p2 = a2 * b2, p3 = a3 * b3;
  return ((p0 + p1) + p2) + p3;
}

// Host-precomputed 16-coefficient cell polynomial.
C8MI_HD double bicubic_precomputed(Tables const& t, c8mi_native_bicubic const& s,
                                   EnergyRow const& er, double loss) {
  if (s.polynomial_coefficient_offset == C8MI_NO_POLYNOMIAL_COEFFICIENTS ||
      t.bicubic_polynomial_coefficients == nullptr)
    return __builtin_nan("");
  double const x1 = axis_transform(s.loss_axis, loss);
  std::uint32_t const n1 = interval_node(x1, s.columns);
  std::uint64_t const cell =
      static_cast<std::uint64_t>(er.n0) + static_cast<std::uint64_t>(s.rows - 1u) * n1;
  std::uint64_t const offset = s.polynomial_coefficient_offset + 16u * cell;
  std::uint64_t const size = t.bicubic_polynomial_coefficient_count;
  if (offset > size || 16u > size - offset) return __builtin_nan("");
  double const* c = t.bicubic_polynomial_coefficients + offset;
  double const f0 = er.f0, f1 = x1 - n1;
  double const v0[4] = {1., f0, f0 * f0, f0 * f0 * f0};
  double const v1[4] = {1., f1, f1 * f1, f1 * f1 * f1};
  double p[4];
#pragma unroll
  for (int row = 0; row < 4; ++row)
    p[row] = dot4(c[4 * row], c[4 * row + 1], c[4 * row + 2], c[4 * row + 3], v1[0], v1[1],
                  v1[2], v1[3]);
  return dot4(v0[0], v0[1], v0[2], v0[3], p[0], p[1], p[2], p[3]);
}

C8MI_HD double bicubic76(Tables const& t, c8mi_native_bicubic const& s, EnergyRow const& er,
                         double loss) {
  if (s.polynomial_coefficient_offset != C8MI_NO_POLYNOMIAL_COEFFICIENTS &&
      t.bicubic_polynomial_coefficients != nullptr)
    return bicubic_precomputed(t, s, er, loss);
  return bicubic_sum(t, s, er, loss, false);
}

C8MI_HD double bicubic76(Tables const& t, c8mi_native_bicubic const& s, double e,
                         double loss) {
  return bicubic76(t, s, energy_row(s, e), loss);
}

// CubicInterpolation 0.1.5's prime(): Boost's sixth-order finite difference.
// The step is (epsilon / 168)^(1/7) rounded to double, 0x1.6deb70960cf79p-9.
C8MI_HD double bicubic_loss_derivative76(Tables const& t, c8mi_native_bicubic const& s,
                                         EnergyRow const& er, double loss) {
  double h = 0x1.6deb70960cf79p-9;
  double const coordinate = axis_transform(s.loss_axis, loss);
  double const representable = coordinate + h;
  h = representable - coordinate;
  auto at = [&](double c) { return bicubic76(t, s, er, axis_back_transform(s.loss_axis, c)); };
  double const y1 = at(coordinate + h) - at(coordinate - h);
  double const y2 = at(coordinate - 2. * h) - at(coordinate + 2. * h);
  double const y3 = at(coordinate + 3. * h) - at(coordinate - 3. * h);
  double const cd = (y3 + 9. * y2 + 45. * y1) / (60. * h);
  return cd * axis_derivative(s.loss_axis, loss);
}

C8MI_HD double cubic(Tables const& t, c8mi_native_cubic const& s, double x) {
  double const coordinate = axis_transform(s.axis, x);
  std::uint32_t const node = interval_node(coordinate, s.coefficient_count);
  double const fraction = coordinate - node;
  std::uint64_t const off = s.coefficient_offset + node;
  double b[4];
  hermite(fraction, b);
  double value = b[0] * t.cubic_values[off] + b[1] * t.cubic_values[off + 1] +
                 b[2] * t.cubic_node_derivatives[off] + b[3] * t.cubic_node_derivatives[off + 1];
  if (s.has_value_transform) value = axis_back_transform(s.value_transform, value);
  return value;
}

C8MI_HD double cubic_derivative76(Tables const& t, c8mi_native_cubic const& s,
                         double x) {
  double const coordinate = axis_transform(s.axis, x);
  std::uint32_t const node = interval_node(coordinate, s.coefficient_count);
  double const fraction = coordinate - node;
  std::uint64_t const off = s.coefficient_offset + node;
  double b[4], d[4];
  hermite(fraction, b);
  hermite_derivative(fraction, d);
  double const raw = b[0] * t.cubic_values[off] + b[1] * t.cubic_values[off + 1] +
                     b[2] * t.cubic_node_derivatives[off] +
                     b[3] * t.cubic_node_derivatives[off + 1];
  double derivative = d[0] * t.cubic_values[off] + d[1] * t.cubic_values[off + 1] +
                      d[2] * t.cubic_node_derivatives[off] +
                      d[3] * t.cubic_node_derivatives[off + 1];
  if (s.has_value_transform) derivative *= axis_back_derivative(s.value_transform, raw);
  return derivative * axis_derivative(s.axis, x);
}

C8MI_HD int sgn(double x) { return (0. < x ? 1 : 0) - (x < 0. ? 1 : 0); }

struct NewtonResult {
  double value;
  std::uint32_t iterations;
  bool runtime_error;
};

// Boost 1.79 newton_raphson_iterate(f, guess, lower, upper, digits = 20).
template <class F, class D>
C8MI_HD NewtonResult newton76(F const& f, D const& df, double initial, double lower,
                     double upper) {
  double residual = 0., last_residual = 0., result = initial, guess = initial;
  double delta = 1.7976931348623157e308, delta1 = delta, delta2 = delta;
  double maximum_residual = 0., minimum_residual = 0.;
  constexpr double factor = 1.9073486328125e-6;  // 2^(1 - 20)
  std::uint32_t it = 0;
  while (it < 256) {
    last_residual = residual;
    delta2 = delta1;
    delta1 = delta;
    residual = f(result);
    double const derivative = df(result);
    ++it;
    if (!isfinite(residual) || !isfinite(derivative)) return {result, it, true};
    if (residual == 0.) break;
    if (derivative == 0.) {
      if (last_residual == 0.) {
        guess = result == lower ? upper : lower;
        last_residual = f(guess);
        delta = guess - result;
      }
      if (sgn(last_residual) * sgn(residual) < 0)
        delta = delta < 0. ? (result - lower) / 2. : (result - upper) / 2.;
      else
        delta = delta < 0. ? (result - upper) / 2. : (result - lower) / 2.;
    } else {
      delta = residual / derivative;
    }
    double const abs_delta2 = delta2 < 0. ? -delta2 : delta2;
    if ((delta * 2. < 0. ? -(delta * 2.) : delta * 2.) > abs_delta2) {
      double const shift = delta > 0. ? (result - lower) / 2. : (result - upper) / 2.;
      double const abs_shift = shift < 0. ? -shift : shift;
      double const abs_result = result < 0. ? -result : result;
      if (result != 0. && abs_shift > abs_result) delta = (delta > 0. ? 1. : -1.) * abs_result;
      else delta = shift;
      delta1 = 3. * delta;
      delta2 = 3. * delta;
    }
    guess = result;
    result -= delta;
    if (result <= lower) {
      delta = .5 * (guess - lower);
      result = guess - delta;
      if (result == lower || result == upper) break;
    } else if (result >= upper) {
      delta = .5 * (guess - upper);
      result = guess - delta;
      if (result == lower || result == upper) break;
    }
    if (delta > 0.) {
      upper = guess;
      maximum_residual = residual;
    } else {
      lower = guess;
      minimum_residual = residual;
    }
    if (maximum_residual * minimum_residual > 0.) return {result, it, true};
    double const abs_result = result < 0. ? -result : result;
    double const abs_step = delta < 0. ? -delta : delta;
    if (!(abs_result * factor < abs_step)) break;
  }
  return {result, it, it == 256};
}

C8MI_HD bool valid_view(Tables const& t) {
  return t.dndx_column_count > 0 && t.dndx_columns != nullptr && t.total_rate_column_count > 0 &&
         t.total_rate_columns != nullptr && t.selection_column_indices != nullptr &&
         (t.dedx_column_count == 0 ||
          (t.dedx_columns != nullptr && t.dedx_accumulation_indices != nullptr)) &&
         t.bicubic_values != nullptr && t.bicubic_derivative_energy != nullptr &&
         t.bicubic_derivative_loss != nullptr && t.bicubic_mixed_derivative != nullptr &&
         t.bicubic_coefficient_count > 0;
}

C8MI_HD c8mi_native_dndx_column const* find_dndx(Tables const& t, std::int32_t pid,
                                        std::int32_t process,
                                        std::uint64_t component) {
  std::uint32_t lo = 0, hi = t.dndx_column_count;
  while (lo < hi) {
    std::uint32_t const mid = lo + (hi - lo) / 2;
    c8mi_native_dndx_column const& c = t.dndx_columns[mid];
    bool const before =
        c.pdg_id < pid ||
        (c.pdg_id == pid && (c.process_id < process ||
                             (c.process_id == process && c.component_hash < component)));
    if (before) lo = mid + 1;
    else hi = mid;
  }
  if (lo < t.dndx_column_count) {
    c8mi_native_dndx_column const& c = t.dndx_columns[lo];
    if (c.pdg_id == pid && c.process_id == process && c.component_hash == component) return &c;
  }
  return nullptr;
}

C8MI_HD void dndx_range(Tables const& t, std::int32_t pid, std::uint32_t& begin,
               std::uint32_t& end) {
  std::uint32_t lo = 0, hi = t.dndx_column_count;
  while (lo < hi) {
    std::uint32_t const mid = lo + (hi - lo) / 2;
    if (t.dndx_columns[mid].pdg_id < pid) lo = mid + 1;
    else hi = mid;
  }
  begin = lo;
  hi = t.dndx_column_count;
  while (lo < hi) {
    std::uint32_t const mid = lo + (hi - lo) / 2;
    if (t.dndx_columns[mid].pdg_id <= pid) lo = mid + 1;
    else hi = mid;
  }
  end = lo;
}

C8MI_HD c8mi_native_dndx_column const* selection_dndx(Tables const& t,
                                             std::uint32_t position) {
  if (position >= t.dndx_column_count) return nullptr;
  std::uint32_t const index = t.selection_column_indices[position];
  if (index >= t.dndx_column_count) return nullptr;
  return &t.dndx_columns[index];
}

C8MI_HD c8mi_native_total_rate_column const* find_total_rate(Tables const& t,
                                                   std::int32_t pid) {
  std::uint32_t lo = 0, hi = t.total_rate_column_count;
  while (lo < hi) {
    std::uint32_t const mid = lo + (hi - lo) / 2;
    if (t.total_rate_columns[mid].pdg_id < pid) lo = mid + 1;
    else hi = mid;
  }
  if (lo < t.total_rate_column_count && t.total_rate_columns[lo].pdg_id == pid)
    return &t.total_rate_columns[lo];
  return nullptr;
}

C8MI_HD c8mi_native_utility_column const* find_utility(Tables const& t,
                                              std::int32_t pid) {
  for (std::uint32_t i = 0; i < t.utility_column_count; ++i)
    if (t.utility_columns[i].pdg_id == pid) return &t.utility_columns[i];
  return nullptr;
}

struct Limits {
  double minimum, maximum;
  bool valid;
};

C8MI_HD Limits kinematic_limits(Tables const& t, c8mi_native_dndx_column const& c,
                       double e) {
  if (!(e > 0.)) return {0., 0., false};
  double const mass = c.particle_mass_mev;
  double minimum = 0., maximum = 0.;
  c8mi_native_constants const& k = t.constants;
  switch (c.kinematic_model) {
    case C8MI_KIN_ONLY_STOCHASTIC:
    case C8MI_KIN_PHOTON_PAIR_FULL_LOSS:
      return {1., 1., true};
    case C8MI_KIN_COMPTON:
      maximum = 1. - 1. / (1. + 2. * e / k.electron_mass_mev);
      break;
    case C8MI_KIN_BREMS_ELECTRON_SCREENING:
      if (!(mass > 0.)) return {0., 0., false};
      maximum = 1. - mass / e;
      if (maximum < 0.) maximum = 0.;
      break;
    case C8MI_KIN_BREMS_KKP:
      if (!(mass > 0.) || !(c.component_nuclear_charge > 0.)) return {0., 0., false};
      maximum = 1. - 0.75 * k.sqrt_e * (mass / e) * pow(c.component_nuclear_charge, 1. / 3.);
      if (maximum < 0.) maximum = 0.;
      break;
    case C8MI_KIN_ELECTRON_PAIR: {
      if (!(mass > 0.)) return {0., 0., false};
      double const ratio = mass / e;
      minimum = 4. * k.electron_mass_mev / e;
      maximum = 1. - 0.75 * k.sqrt_e * ratio * pow(c.component_nuclear_charge, 1. / 3.);
      double const second = 1. - 6. * ratio * ratio;
      if (second < maximum) maximum = second;
      if (maximum < minimum) maximum = minimum;
      break;
    }
    case C8MI_KIN_PHOTONUCLEAR_ALLM97: {
      double const nucleon = c.component_average_nucleon_mass_mev, pion = k.charged_pion_mass_mev;
      if (!(nucleon > 0.) || !(pion > 0.) || !(mass > 0.)) return {0., 0., false};
      minimum = (pion + pion * pion / (2. * nucleon)) / e;
      maximum = 1.;
      if (mass < pion) {
        double const ratio = mass / nucleon;
        maximum -= nucleon * (1. + ratio * ratio) / (2. * e);
      }
      if (maximum < minimum) maximum = minimum;
      break;
    }
    case C8MI_KIN_IONIZATION_MOLLER:
      if (!(mass > 0.)) return {0., 0., false};
      maximum = 0.5 * (1. - mass / e);
      if (maximum < 0.) maximum = 0.;
      break;
    case C8MI_KIN_IONIZATION_BHABHA:
      if (!(mass > 0.)) return {0., 0., false};
      maximum = 1. - mass / e;
      if (maximum < 0.) maximum = 0.;
      break;
    case C8MI_KIN_IONIZATION_BETHE_BLOCH_ROSSI: {
      if (!(mass > 0.) || !(c.medium_mean_excitation_energy_mev > 0.)) return {0., 0., false};
      double const mass_ratio = k.electron_mass_mev / mass, gamma = e / mass;
      minimum = c.medium_mean_excitation_energy_mev / e;
      maximum = 2. * k.electron_mass_mev * (gamma * gamma - 1.) /
                ((1. + 2. * gamma * mass_ratio + mass_ratio * mass_ratio) * e);
      double const kinetic = 1. - mass / e;
      if (kinetic < maximum) maximum = kinetic;
      if (maximum < minimum) maximum = minimum;
      break;
    }
    default:
      return {0., 0., false};
  }
  if (!isfinite(minimum) || !isfinite(maximum) || maximum < minimum) return {0., 0., false};
  double cut = minimum;
  if (c.absolute_cut_mev > 0. && c.relative_cut > 0.) {
    double requested = c.absolute_cut_mev / e;
    if (c.relative_cut < requested) requested = c.relative_cut;
    if (requested < minimum) requested = minimum;
    if (requested > maximum) requested = maximum;
    cut = requested;
  }
  return {cut, maximum, true};
}

C8MI_HD double physical_loss(Tables const& t, c8mi_native_dndx_column const& c,
                    double e, double loss) {
  Limits const l = kinematic_limits(t, c, e);
  if (!l.valid) return __builtin_nan("");
  double const lower = l.minimum, upper = l.maximum;
  if (loss <= 0.) return lower;
  if (loss >= 1.) return upper;
  if (lower == upper) return lower;
  switch (c.loss_transform) {
    case C8MI_LOSS_RELATIVE_LOG:
      if (!(lower > 0.)) return __builtin_nan("");
      return lower * exp(loss * log(upper / lower));
    case C8MI_LOSS_RELATIVE_LOG_POW15:
      if (!(lower > 0.)) return __builtin_nan("");
      return lower * exp(pow(loss, 1.5) * log(upper / lower));
    case C8MI_LOSS_ONE_MINUS_LOG: {
      double const xi = log((1. - lower) / (1. - upper));
      return 1. - (1. - upper) * exp((1. - loss) * xi);
    }
  }
  return __builtin_nan("");
}

C8MI_HD bool rate_below_threshold(c8mi_native_dndx_column const& c, double e) {
  return c.rate_model == C8MI_RATE_BICUBIC &&
         (e < c.lower_energy_limit_mev || e < c.spline.energy_axis.low);
}

}  // namespace detail

C8MI_HD QueryResult cumulative_rate(Tables const& t, c8mi_native_dndx_column const& c,
                                    double e, double loss) {
  using namespace detail;
  if (!valid_view(t)) return qr(C8MI_QUERY_INVALID_VIEW);
  if (!isfinite(e) || !isfinite(loss)) return qr(C8MI_QUERY_NON_FINITE_INPUT);
  if (e < c.lower_energy_limit_mev || e < c.spline.energy_axis.low) return qr(C8MI_QUERY_SUCCESS);
  if (e > c.spline.energy_axis.high) return qr(C8MI_QUERY_ENERGY_OUT_OF_RANGE);
  if (loss < c.spline.loss_axis.low || loss > c.spline.loss_axis.high)
    return qr(C8MI_QUERY_QUANTILE_OUT_OF_RANGE);
  double value = bicubic76(t, c.spline, e, loss) / c.component_weight;
  if (value < 0.) value = 0.;
  if (!isfinite(value)) return qr(C8MI_QUERY_INVALID_VIEW);
  return qr(C8MI_QUERY_SUCCESS, value);
}

C8MI_HD QueryResult rate_for_column(Tables const& t, c8mi_native_dndx_column const& c,
                                    double e) {
  using namespace detail;
  if (!valid_view(t)) return qr(C8MI_QUERY_INVALID_VIEW);
  if (rate_below_threshold(c, e)) return qr(C8MI_QUERY_SUCCESS);
  if (c.rate_model == C8MI_RATE_ANNIHILATION_HEITLER) {
    if (!isfinite(e)) return qr(C8MI_QUERY_NON_FINITE_INPUT);
    if (e <= c.particle_mass_mev) return qr(C8MI_QUERY_SUCCESS);
    double const gamma = e / c.particle_mass_mev;
    double const root = pow(gamma * gamma - 1., .5);
    double const aux = (gamma * gamma + 4. * gamma + 1.) / (gamma * gamma - 1.) * log(gamma + root) -
                       (gamma + 3.) / root;
    return qr(C8MI_QUERY_SUCCESS, aux * c.direct_rate_factor / (gamma + 1.));
  }
  if (c.rate_model == C8MI_RATE_PHOTOEFFECT_SAUTER) {
    if (!isfinite(e)) return qr(C8MI_QUERY_NON_FINITE_INPUT);
    if (!(e > c.lower_energy_limit_mev)) return qr(C8MI_QUERY_SUCCESS);
    double const alpha = t.constants.fine_structure_constant, me = t.constants.electron_mass_mev;
    double const pi = t.constants.pi, re = t.constants.classical_electron_radius_cm;
    double const z = c.component_nuclear_charge;
    double const ionization = z * z * alpha * alpha * me / 2.;
    double const gamma = 1. + (e - ionization) / me;
    double const gamma_beta = sqrt(gamma * gamma - 1.);
    double const sigma_t = 8. * pi * re * re / 3.;
    double const sigma = 1.5 * sigma_t * pow(z, 5.) * pow(alpha, 4.) * pow(me / e, 5.) *
                         pow(gamma_beta, 3.) *
                         (4. / 3. + gamma * (gamma - 2.) / (gamma + 1.) *
                                        (1. - log(gamma + gamma_beta) / (gamma * gamma_beta)));
    double const zab = z * alpha * gamma / gamma_beta;
    double const zpab = pi * zab;
    double correction;
    if (zpab < 10.)
      correction = (1. + zab * zab) * zpab / sinh(zpab) * exp(zab * (pi - 4. * atan(1. / zab)));
    else
      correction = (1. + zab * zab) * 2. * zpab * exp(-4. * zab * atan(1. / zab));
    double const log_z = log(z);
    double const shell = 1. + 0.01481 * log_z * log_z - 0.000788 * log_z * log_z * log_z;
    return qr(C8MI_QUERY_SUCCESS, sigma * correction * shell * c.direct_rate_factor);
  }
  if (c.rate_model == C8MI_RATE_PHOTOPRODUCTION_HECK) {
    if (!isfinite(e)) return qr(C8MI_QUERY_NON_FINITE_INPUT);
    if (e < c.lower_energy_limit_mev) return qr(C8MI_QUERY_SUCCESS);
    double const eg = e * 1.e-3, mn = c.component_average_nucleon_mass_mev * 1.e-3;
    double const s = mn * mn + 2. * mn * eg;
    constexpr double s0 = 1.0761 * 1.0761;
    double cross;
    if (sqrt(s) < 19.39)
      cross = (73.7 * pow(s, 0.073) + 191.7 * pow(s, -0.602)) * sqrt(1. - s0 / s);
    else
      cross = 59.3 * pow(s, 0.093) + 120.2 * pow(s, -0.358);
    constexpr double rm[3] = {1.231, 1.515, 1.680}, rw[3] = {0.11, 0.11, 0.125};
    constexpr double rs[3] = {31.125, 25.567, 17.508}, rwin[3] = {0.17, 0.38, 0.38};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double const m2 = rm[i] * rm[i], w2 = rw[i] * rw[i];
      double const bw = s / (eg * eg) * rs[i] * w2 * s / ((s - m2) * (s - m2) + w2 * s);
      double th = 0.;
      if (eg >= 0.152 + rwin[i]) th = 1.;
      else if (eg >= 0.152) th = (eg - 0.152) / rwin[i];
      cross += bw * th;
    }
    return qr(C8MI_QUERY_SUCCESS, cross * pow(c.component_atomic_mass, 0.91) * c.direct_rate_factor);
  }
  if (c.rate_model == C8MI_RATE_BICUBIC && (c.pdg_id == 13 || c.pdg_id == -13)) {
    if (!isfinite(e)) return qr(C8MI_QUERY_NON_FINITE_INPUT);
    if (e > c.spline.energy_axis.high) return qr(C8MI_QUERY_ENERGY_OUT_OF_RANGE);
    double value = bicubic_precomputed(t, c.spline, energy_row(c.spline, e), 1.) / c.component_weight;
    if (value < 0.) value = 0.;
    if (!isfinite(value)) return qr(C8MI_QUERY_INVALID_VIEW);
    return qr(C8MI_QUERY_SUCCESS, value);
  }
  return cumulative_rate(t, c, e, 1.);
}

C8MI_HD QueryResult query_rate(Tables const& t, std::int32_t pid, std::int32_t process,
                      std::uint64_t component, double e) {
  using namespace detail;
  if (!valid_view(t)) return qr(C8MI_QUERY_INVALID_VIEW);
  auto const* c = find_dndx(t, pid, process, component);
  if (!c) return qr(C8MI_QUERY_COLUMN_NOT_FOUND);
  if (rate_below_threshold(*c, e)) return qr(C8MI_QUERY_SUCCESS);
  return rate_for_column(t, *c, e);
}

C8MI_HD QueryResult total_rate(Tables const& t, std::int32_t pid, double e) {
  using namespace detail;
  if (!valid_view(t)) return qr(C8MI_QUERY_INVALID_VIEW);
  if (!isfinite(e)) return qr(C8MI_QUERY_NON_FINITE_INPUT);
  auto const* c = find_total_rate(t, pid);
  if (!c) return qr(C8MI_QUERY_PARTICLE_NOT_FOUND);
  if (e < c->lower_energy_limit_mev) return qr(C8MI_QUERY_SUCCESS);
  if (e < c->spline.axis.low || e > c->spline.axis.high) return qr(C8MI_QUERY_ENERGY_OUT_OF_RANGE);
  double const value = cubic(t, c->spline, e);
  if (!isfinite(value)) return qr(C8MI_QUERY_INVALID_VIEW);
  return qr(C8MI_QUERY_SUCCESS, value < 0. ? 0. : value);
}

C8MI_HD QueryResult selection_total_rate(Tables const& t, std::int32_t pid, double e) {
  using namespace detail;
  if (!valid_view(t)) return qr(C8MI_QUERY_INVALID_VIEW);
  double total = 0.;
  std::uint32_t begin, end;
  dndx_range(t, pid, begin, end);
  for (std::uint32_t i = begin; i < end; ++i) {
    auto const* c = selection_dndx(t, i);
    if (!c || c->pdg_id != pid) return qr(C8MI_QUERY_INVALID_VIEW);
    if (rate_below_threshold(*c, e)) continue;
    QueryResult const r = rate_for_column(t, *c, e);
    if (r.status != C8MI_QUERY_SUCCESS) return r;
    total += r.value;
  }
  if (begin == end) return qr(C8MI_QUERY_PARTICLE_NOT_FOUND);
  return qr(C8MI_QUERY_SUCCESS, total);
}

C8MI_HD QueryResult dedx(Tables const& t, std::int32_t pid, double e) {
  using namespace detail;
  if (!isfinite(e)) return qr(C8MI_QUERY_NON_FINITE_INPUT);
  double total = 0.;
  bool found = false;
  for (std::uint32_t p = 0; p < t.dedx_column_count; ++p) {
    std::uint32_t const index = t.dedx_accumulation_indices[p];
    if (index >= t.dedx_column_count) return qr(C8MI_QUERY_INVALID_VIEW);
    c8mi_native_dedx_column const& c = t.dedx_columns[index];
    if (c.pdg_id != pid) continue;
    found = true;
    if (e < c.lower_energy_limit_mev) continue;
    if (e < c.spline.axis.low || e > c.spline.axis.high) return qr(C8MI_QUERY_ENERGY_OUT_OF_RANGE);
    total += cubic(t, c.spline, e) / c.component_weight;
  }
  if (!found) return qr(C8MI_QUERY_PARTICLE_NOT_FOUND);
  return qr(C8MI_QUERY_SUCCESS, total);
}

C8MI_HD QueryResult loss_fraction(Tables const& t, std::int32_t pid, std::int32_t process,
                         std::uint64_t component, double e, double q,
                         bool photon_split) {
  using namespace detail;
  if (!valid_view(t)) return qr(C8MI_QUERY_INVALID_VIEW);
  if (!isfinite(e) || !isfinite(q)) return qr(C8MI_QUERY_NON_FINITE_INPUT);
  if (q < 0. || q > 1.) return qr(C8MI_QUERY_QUANTILE_OUT_OF_RANGE);
  auto const* cp = find_dndx(t, pid, process, component);
  if (!cp) return qr(C8MI_QUERY_COLUMN_NOT_FOUND);
  c8mi_native_dndx_column const& c = *cp;
  bool const muon_split =
      photon_split && process == 1000000015 && c.kinematic_model == C8MI_KIN_ONLY_STOCHASTIC;
  if (photon_split &&
      (pid != 22 || (!muon_split && c.kinematic_model != C8MI_KIN_PHOTON_PAIR_FULL_LOSS) ||
       c.loss_transform != C8MI_LOSS_RELATIVE_LOG))
    return qr(C8MI_QUERY_UNSUPPORTED_KINEMATICS);
  if ((c.kinematic_model == C8MI_KIN_ONLY_STOCHASTIC && !muon_split) ||
      (c.kinematic_model == C8MI_KIN_PHOTON_PAIR_FULL_LOSS && !photon_split))
    return qr(C8MI_QUERY_SUCCESS, 1.);
  if (c.rate_model != C8MI_RATE_BICUBIC) return qr(C8MI_QUERY_UNSUPPORTED_KINEMATICS);
  double xmin = t.constants.electron_mass_mev / e;
  double xmax = 1. - xmin;
  if (muon_split) {
    double const threshold = 2. * t.constants.muon_mass_mev * t.constants.sqrt_e;
    if (!(t.constants.muon_mass_mev > 0.) || !(e > threshold))
      return qr(C8MI_QUERY_ENERGY_OUT_OF_RANGE);
    double const width = sqrt(1. - 2. * t.constants.sqrt_e * t.constants.muon_mass_mev / e);
    xmin = .5 * (1. - width);
    xmax = .5 * (1. + width);
  }
  if (photon_split &&
      (xmin >= .5 || e < c.spline.energy_axis.low || e > c.spline.energy_axis.high))
    return qr(C8MI_QUERY_ENERGY_OUT_OF_RANGE);
  Limits const lim = photon_split ? Limits{xmin, xmax, true} : kinematic_limits(t, c, e);
  if (!lim.valid) return qr(C8MI_QUERY_UNSUPPORTED_KINEMATICS);
  if (lim.minimum == lim.maximum) return qr(C8MI_QUERY_SUCCESS, lim.minimum);
  QueryResult total;
  if (!photon_split) {
    total = cumulative_rate(t, c, e, 1.);
  } else {
    double const raw = bicubic76(t, c.spline, e, 1.);
    total = isfinite(raw) ? qr(C8MI_QUERY_SUCCESS, raw < 0. ? 0. : raw) : qr(C8MI_QUERY_INVALID_VIEW);
  }
  if (total.status != C8MI_QUERY_SUCCESS) return total;
  double const target = photon_split ? q * total.value : (q * total.value) * c.component_weight;
  if (q == 0.) return qr(C8MI_QUERY_SUCCESS, lim.minimum);
  if (q == 1.) return qr(C8MI_QUERY_SUCCESS, lim.maximum);
  EnergyRow const row = energy_row(c.spline, e);
  auto residual_at = [&](double loss) { return bicubic76(t, c.spline, row, loss) - target; };

  double lower = c.spline.loss_axis.low, upper = c.spline.loss_axis.high;
  double lower_residual = residual_at(lower);
  double upper_residual = residual_at(upper);
  std::uint32_t bisection = 0;
  bool need_bisection = false, exact = false;
  double current = 0.;
  if (!isfinite(lower_residual) || !isfinite(upper_residual)) {
    return qr(C8MI_QUERY_ROOT_NOT_CONVERGED);
  } else if (lower_residual == 0.) {
    current = lower;
    exact = true;
  } else if (upper_residual == 0.) {
    current = upper;
    exact = true;
  } else if (lower_residual * upper_residual >= 0.) {
    need_bisection = true;
  } else {
    double const coarse = (upper - lower) * 1.e-2;
    while (!((upper - lower) < coarse)) {
      double const middle = (lower + upper) / 2.;
      double const middle_residual = residual_at(middle);
      ++bisection;
      if (!isfinite(middle_residual)) {
        need_bisection = true;
        break;
      }
      if (middle == upper || middle == lower) break;
      if (middle_residual == 0.) {
        lower = upper = middle;
        exact = true;
        break;
      }
      if (sgn(middle_residual) * sgn(lower_residual) < 0) {
        upper = middle;
        upper_residual = middle_residual;
      } else {
        lower = middle;
        lower_residual = middle_residual;
      }
    }
    current = (lower + upper) / 2.;
  }

  double delta = 1.7976931348623157e308, delta1 = delta, delta2 = delta;
  double previous_residual = 0., maximum_residual = 0., minimum_residual = 0.;
  constexpr double factor = 1.9073486328125e-6;
  std::uint32_t newton = 0;
  double guess = current;
  while (!exact && !need_bisection && newton < 256) {
    double const last_residual = previous_residual;
    delta2 = delta1;
    delta1 = delta;
    double const residual = residual_at(current);
    previous_residual = residual;
    ++newton;
    if (!isfinite(residual)) {
      need_bisection = true;
      break;
    }
    if (residual == 0.) {
      exact = true;
      break;
    }
    double const derivative = bicubic_loss_derivative76(t, c.spline, row, current);
    if (!isfinite(derivative)) {
      need_bisection = true;
      break;
    } else if (derivative == 0.) {
      double mutable_last = last_residual;
      if (mutable_last == 0.) {
        guess = current == lower ? upper : lower;
        mutable_last = residual_at(guess);
        delta = guess - current;
      }
      if (sgn(mutable_last) * sgn(residual) < 0)
        delta = delta < 0. ? (current - lower) / 2. : (current - upper) / 2.;
      else
        delta = delta < 0. ? (current - upper) / 2. : (current - lower) / 2.;
    } else {
      delta = residual / derivative;
    }
    double const abs_delta = delta < 0. ? -delta : delta;
    double const abs_delta2 = delta2 < 0. ? -delta2 : delta2;
    if (2. * abs_delta > abs_delta2) {
      double const shift = delta > 0. ? (current - lower) * .5 : (current - upper) * .5;
      double const abs_shift = shift < 0. ? -shift : shift;
      double const abs_current = current < 0. ? -current : current;
      if (current != 0. && abs_shift > abs_current) delta = (delta > 0. ? 1. : -1.) * abs_current;
      else delta = shift;
      delta1 = 3. * delta;
      delta2 = 3. * delta;
    }
    guess = current;
    current -= delta;
    if (current <= lower) {
      delta = .5 * (guess - lower);
      current = guess - delta;
      if (current == lower || current == upper) break;
    } else if (current >= upper) {
      delta = .5 * (guess - upper);
      current = guess - delta;
      if (current == lower || current == upper) break;
    }
    if (delta > 0.) {
      upper = guess;
      maximum_residual = residual;
    } else {
      lower = guess;
      minimum_residual = residual;
    }
    if (maximum_residual * minimum_residual > 0.) {
      need_bisection = true;
      break;
    }
    double const abs_current = current < 0. ? -current : current;
    double const abs_step = delta < 0. ? -delta : delta;
    if (!(abs_current * factor < abs_step)) break;
  }
  if (newton == 256) need_bisection = true;

  if (need_bisection) {
    lower = c.spline.loss_axis.low;
    upper = c.spline.loss_axis.high;
    lower_residual = residual_at(lower);
    upper_residual = residual_at(upper);
    if (!isfinite(lower_residual) || !isfinite(upper_residual) ||
        lower_residual * upper_residual > 0.)
      return qr(C8MI_QUERY_ROOT_NOT_CONVERGED, 0., newton, bisection);
    for (std::uint32_t it = 0; it <= 100; ++it) {
      double const middle = (lower + upper) / 2.;
      double const middle_residual = residual_at(middle);
      double const current_lower = residual_at(lower);
      if (!isfinite(middle_residual) || !isfinite(current_lower))
        return qr(C8MI_QUERY_ROOT_NOT_CONVERGED, 0., newton, bisection);
      if (sgn(middle_residual) == sgn(current_lower)) lower = middle;
      else upper = middle;
      ++bisection;
      if ((upper - lower) < 1.e-6) break;
    }
    current = (lower + upper) / 2.;
  }
  double const physical = photon_split
                              ? lim.minimum * exp(current * log(lim.maximum / lim.minimum))
                              : physical_loss(t, c, e, current);
  if (!isfinite(physical)) return qr(C8MI_QUERY_ROOT_NOT_CONVERGED, 0., newton, bisection);
  return qr(C8MI_QUERY_SUCCESS, physical, newton, bisection);
}

C8MI_HD QueryResult range(Tables const& t, std::int32_t pid, double e) {
  using namespace detail;
  auto const* c = find_utility(t, pid);
  if (!c) return qr(C8MI_QUERY_PARTICLE_NOT_FOUND);
  if (!isfinite(e)) return qr(C8MI_QUERY_NON_FINITE_INPUT);
  if (e < c->lower_energy_limit_mev || e < c->spline.axis.low || e > c->spline.axis.high)
    return qr(C8MI_QUERY_ENERGY_OUT_OF_RANGE);
  double const upper = cubic(t, c->spline, e);
  double const lower = cubic(t, c->spline, c->lower_energy_limit_mev);
  double const value = c->reverse ? lower - upper : upper - lower;
  if (!isfinite(value) || value < 0.) return qr(C8MI_QUERY_INVALID_VIEW);
  return qr(C8MI_QUERY_SUCCESS, value);
}

C8MI_HD QueryResult energy_after_loss(Tables const& t, std::int32_t pid, double e0,
                             double grammage) {
  using namespace detail;
  auto const* c = find_utility(t, pid);
  if (!c) return qr(C8MI_QUERY_PARTICLE_NOT_FOUND);
  if (!isfinite(e0) || !isfinite(grammage) || grammage < 0.) return qr(C8MI_QUERY_NON_FINITE_INPUT);
  if (e0 < c->lower_energy_limit_mev || e0 > c->spline.axis.high)
    return qr(C8MI_QUERY_ENERGY_OUT_OF_RANGE);
  QueryResult const maximum = range(t, pid, e0);
  if (maximum.status != C8MI_QUERY_SUCCESS) return maximum;
  if (grammage > maximum.value) return qr(C8MI_QUERY_ENERGY_OUT_OF_RANGE);
  if (grammage == maximum.value) return qr(C8MI_QUERY_SUCCESS, c->lower_energy_limit_mev);
  if (grammage == 0.) return qr(C8MI_QUERY_SUCCESS, e0);
  double const signed_grammage = c->reverse ? -grammage : grammage;
  double const target = cubic(t, c->spline, e0) - signed_grammage;
  auto residual = [&](double e) { return cubic(t, c->spline, e) - target; };
  auto slope = [&](double e) { return cubic_derivative76(t, c->spline, e); };
  auto bisect = [&](double accuracy, double& lower, double& upper, std::uint32_t& iterations) {
    double const lv = residual(lower), uv = residual(upper);
    if (!isfinite(lv) || !isfinite(uv) || lv * uv > 0.) return false;
    for (std::uint32_t it = 0; it <= 100; ++it) {
      double const center = (lower + upper) / 2.;
      double const cv = residual(center);
      double const clv = residual(lower);
      if (!isfinite(cv) || !isfinite(clv)) return false;
      if (sgn(cv) == sgn(clv)) lower = center;
      else upper = center;
      ++iterations;
      if ((upper - lower) < accuracy) return true;
    }
    return true;
  };
  double lower = c->lower_energy_limit_mev, upper = e0;
  std::uint32_t bisection = 0;
  double const coarse = (e0 - c->lower_energy_limit_mev) * 1.e-2;
  if (!bisect(coarse, lower, upper, bisection))
    return qr(C8MI_QUERY_ROOT_NOT_CONVERGED, 0., 0, bisection);
  double const initial = upper == e0 ? e0 : (lower + upper) / 2.;
  NewtonResult const n = newton76(residual, slope, initial, lower, upper);
  if (!n.runtime_error && isfinite(n.value))
    return qr(C8MI_QUERY_SUCCESS, n.value, n.iterations, bisection);
  lower = c->lower_energy_limit_mev;
  upper = e0;
  if (!bisect(1.e-6, lower, upper, bisection))
    return qr(C8MI_QUERY_ROOT_NOT_CONVERGED, 0., n.iterations, bisection);
  return qr(C8MI_QUERY_SUCCESS, lower, n.iterations, bisection);
}

C8MI_HD QueryResult energy_of_range(Tables const& t, std::int32_t pid, double r) {
  using namespace detail;
  auto const* c = find_utility(t, pid);
  if (!c) return qr(C8MI_QUERY_PARTICLE_NOT_FOUND);
  if (!isfinite(r) || r < 0.) return qr(C8MI_QUERY_NON_FINITE_INPUT);
  QueryResult const maximum = range(t, pid, c->spline.axis.high);
  if (maximum.status != C8MI_QUERY_SUCCESS) return maximum;
  if (r > maximum.value) return qr(C8MI_QUERY_ENERGY_OUT_OF_RANGE);
  double lower = c->lower_energy_limit_mev, upper = c->spline.axis.high;
  std::uint32_t it = 0;
  for (; it < 100; ++it) {
    double const lc = axis_transform(c->spline.axis, lower);
    double const uc = axis_transform(c->spline.axis, upper);
    double const middle = axis_back_transform(c->spline.axis, 0.5 * (lc + uc));
    QueryResult const value = range(t, pid, middle);
    if (value.status != C8MI_QUERY_SUCCESS) return value;
    if (value.value < r) lower = middle;
    else upper = middle;
    if ((upper - lower) <= 1.e-11 * upper) break;
  }
  return qr(C8MI_QUERY_SUCCESS, 0.5 * (lower + upper), 0, it);
}

C8MI_HD QueryResult run_query(Tables const& t, c8mi_native_query const& q) {
  switch (q.kind) {
    case C8MI_NATIVE_TOTAL_RATE: return total_rate(t, q.pdg_id, q.energy_mev);
    case C8MI_NATIVE_SELECTION_TOTAL_RATE: return selection_total_rate(t, q.pdg_id, q.energy_mev);
    case C8MI_NATIVE_RATE:
      return query_rate(t, q.pdg_id, q.process_id, q.component_hash, q.energy_mev);
    case C8MI_NATIVE_DEDX: return dedx(t, q.pdg_id, q.energy_mev);
    case C8MI_NATIVE_LOSS_FRACTION:
    case C8MI_NATIVE_PHOTON_SPLIT:
      return loss_fraction(t, q.pdg_id, q.process_id, q.component_hash, q.energy_mev, q.argument,
                           q.kind == C8MI_NATIVE_PHOTON_SPLIT);
    case C8MI_NATIVE_RANGE: return range(t, q.pdg_id, q.energy_mev);
    case C8MI_NATIVE_ENERGY_AFTER_LOSS:
      return energy_after_loss(t, q.pdg_id, q.energy_mev, q.argument);
    case C8MI_NATIVE_ENERGY: return energy_of_range(t, q.pdg_id, q.argument);
    case C8MI_NATIVE_CUMULATIVE_RATE: {
      auto const* c = detail::find_dndx(t, q.pdg_id, q.process_id, q.component_hash);
      return c ? cumulative_rate(t, *c, q.energy_mev, q.argument) : qr(C8MI_QUERY_COLUMN_NOT_FOUND);
    }
  }
  return qr(C8MI_QUERY_INVALID_VIEW);
}

}  // namespace c8mi::em
