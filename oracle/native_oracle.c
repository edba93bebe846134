/*
 * c8mi oracle — TEST INFRASTRUCTURE ONLY (see c8_oracle.h).
 *
 * FP64 restatement of
 *  - the Philox4x32-10 stream and its key derivation
 *    (corsika/detail/kokkos/em/random/Philox.inl:15-57; Random123's published
 *    Philox4x32-10: multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key increments
 *    0x9E3779B9 / 0xBB67AE85, ten rounds);
 *  - the PROPOSAL-native table queries
 *    (corsika/detail/kokkos/em/tables/ProposalNativeQueries.inl:88-1428).
 * Built without floating-point contraction, like the reference's host build.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "c8_oracle.h"

/* ---------------- Philox (Philox.inl:15-57) ---------------- */

static uint64_t splitmix64(uint64_t v) {
  v += 0x9e3779b97f4a7c15ULL;
  v = (v ^ (v >> 30U)) * 0xbf58476d1ce4e5b9ULL;
  v = (v ^ (v >> 27U)) * 0x94d049bb133111ebULL;
  return v ^ (v >> 31U);
}

void oracle_philox4x32_10(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = counter[0], c1 = counter[1], c2 = counter[2], c3 = counter[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    if (round > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

void oracle_random_words(uint64_t seed, uint64_t shower, uint64_t history, uint64_t step,
                         uint32_t process, uint64_t draw, uint32_t out[4]) {
  uint32_t counter[4] = {(uint32_t)history, (uint32_t)(history >> 32U), (uint32_t)step,
                         (uint32_t)(step >> 32U)};
  uint64_t m = splitmix64(seed);
  m = splitmix64(m ^ shower);
  m = splitmix64(m ^ (uint64_t)process);
  m = splitmix64(m ^ draw);
  uint32_t key[2] = {(uint32_t)m, (uint32_t)(m >> 32U)};
  oracle_philox4x32_10(counter, key, out);
}

double oracle_uniform_open01(uint64_t seed, uint64_t shower, uint64_t history, uint64_t step,
                             uint32_t process, uint64_t draw, uint32_t lane) {
  uint32_t w[4];
  oracle_random_words(seed, shower, history, step, process, draw, w);
  return ((double)w[lane & 3U] + 0.5) * (1. / 4294967296.);
}

/* ---------------- native tables (ProposalNativeQueries.inl) ---------------- */

typedef c8mi_native_table_set view_t;
typedef c8mi_native_result result_t;

static const double kLn2 = 0.693147180559945309417232121458176568;

static result_t res(uint32_t status, uint32_t newton, uint32_t bisection, double value) {
  result_t r;
  memset(&r, 0, sizeof r);
  r.status = status;
  r.newton_iterations = newton;
  r.bisection_iterations = bisection;
  r.value = value;
  return r;
}

/* axis_transform / axis_back_transform / derivatives (:88-144) */
static double axis_transform(const c8mi_native_axis* a, double x) {
  switch (a->type) {
    case C8MI_AXIS_LINEAR: return (x - a->low) / a->step;
    case C8MI_AXIS_EXPONENTIAL: return log(x / a->low) / a->step;
    case C8MI_AXIS_EXPONENTIAL_MINUS_ONE: return (log1p(x / a->low) - kLn2) / a->step;
  }
  return NAN;
}
static double axis_back_transform(const c8mi_native_axis* a, double c) {
  switch (a->type) {
    case C8MI_AXIS_LINEAR: return c * a->step + a->low;
    case C8MI_AXIS_EXPONENTIAL: return a->low * exp(c * a->step);
    case C8MI_AXIS_EXPONENTIAL_MINUS_ONE: return a->low * expm1(c * a->step + kLn2);
  }
  return NAN;
}
static double axis_derivative(const c8mi_native_axis* a, double x) {
  switch (a->type) {
    case C8MI_AXIS_LINEAR: return 1. / a->step;
    case C8MI_AXIS_EXPONENTIAL: return 1. / (x * a->step);
    case C8MI_AXIS_EXPONENTIAL_MINUS_ONE: return 1. / (a->step * (x + a->low));
  }
  return NAN;
}
static double axis_back_derivative(const c8mi_native_axis* a, double c) {
  switch (a->type) {
    case C8MI_AXIS_LINEAR: return a->step;
    case C8MI_AXIS_EXPONENTIAL: return a->low * a->step * exp(c * a->step);
    case C8MI_AXIS_EXPONENTIAL_MINUS_ONE: return a->low * a->step * exp(c * a->step + kLn2);
  }
  return NAN;
}

/* interval_node (:146-153), including the NaN path (-1 cast to uint32). */
static uint32_t interval_node(double coordinate, uint32_t nodes) {
  long long node = (long long)(coordinate >= 0. ? coordinate : -1.);
  if (coordinate < 0.) node = 0;
  if (node > (long long)nodes - 2) node = (long long)nodes - 2;
  return (uint32_t)node;
}

static void hermite(double x, double b[4]) {
  double x2 = x * x, x3 = x2 * x;
  b[0] = 1. - 3. * x2 + 2. * x3;
  b[1] = 3. * x2 - 2. * x3;
  b[2] = x - 2. * x2 + x3;
  b[3] = -x2 + x3;
}
static void hermite_d(double x, double b[4]) {
  double x2 = x * x;
  b[0] = -6. * x + 6. * x2;
  b[1] = 6. * x - 6. * x2;
  b[2] = 1. - 4. * x + 3. * x2;
  b[3] = -2. * x + 3. * x2;
}

static const double* bicubic_array(const view_t* v, uint32_t a, uint32_t b) {
  if (a >= 2 && b < 2) return v->bicubic_derivative_energy;
  if (a < 2 && b >= 2) return v->bicubic_derivative_loss;
  if (a >= 2 && b >= 2) return v->bicubic_mixed_derivative;
  return v->bicubic_values;
}

/* evaluate_bicubic (:181-209) and its loss derivative (:261-299). */
static double bicubic_sum(const view_t* v, const c8mi_native_bicubic* s, double e, double t,
                          int derivative) {
  double x0 = axis_transform(&s->energy_axis, e);
  double x1 = axis_transform(&s->loss_axis, t);
  uint32_t n0 = interval_node(x0, s->rows);
  uint32_t n1 = interval_node(x1, s->columns);
  double h0[4], h1[4];
  hermite(x0 - n0, h0);
  if (derivative)
    hermite_d(x1 - n1, h1);
  else
    hermite(x1 - n1, h1);
  double r = 0.;
  for (uint32_t a = 0; a < 4; ++a) {
    for (uint32_t b = 0; b < 4; ++b) {
      uint32_t row = n0 + (a & 1u), col = n1 + (b & 1u);
      uint64_t local = (uint64_t)row + (uint64_t)s->rows * col;
      r += h0[a] * bicubic_array(v, a, b)[s->coefficient_offset + local] * h1[b];
    }
  }
  return derivative ? r / s->loss_axis.step : r;
}

static double dot4(double a0, double a1, double a2, double a3, double b0, double b1, double b2,
                   double b3) {
  double p0 = a0 * b0, p1 = a1 * b1, p2 = a2 * b2, p3 = a3 * b3;
  return ((p0 + p1) + p2) + p3;
}

/* evaluate_bicubic_precomputed (:222-250) */
static double bicubic_precomputed(const view_t* v, const c8mi_native_bicubic* s, double e,
                                  double t) {
  if (s->polynomial_coefficient_offset == C8MI_NO_POLYNOMIAL_COEFFICIENTS ||
      v->bicubic_polynomial_coefficients == NULL)
    return NAN;
  double x0 = axis_transform(&s->energy_axis, e);
  double x1 = axis_transform(&s->loss_axis, t);
  uint32_t n0 = interval_node(x0, s->rows);
  uint32_t n1 = interval_node(x1, s->columns);
  uint64_t cell = (uint64_t)n0 + (uint64_t)(s->rows - 1u) * n1;
  uint64_t offset = s->polynomial_coefficient_offset + 16u * cell;
  uint64_t size = v->bicubic_polynomial_coefficient_count;
  if (offset > size || 16u > size - offset) return NAN;
  const double* c = v->bicubic_polynomial_coefficients + offset;
  double f0 = x0 - n0, f1 = x1 - n1;
  double v0[4] = {1., f0, f0 * f0, f0 * f0 * f0};
  double v1[4] = {1., f1, f1 * f1, f1 * f1 * f1};
  double p[4];
  for (int row = 0; row < 4; ++row)
    p[row] = dot4(c[4 * row], c[4 * row + 1], c[4 * row + 2], c[4 * row + 3], v1[0], v1[1],
                  v1[2], v1[3]);
  return dot4(v0[0], v0[1], v0[2], v0[3], p[0], p[1], p[2], p[3]);
}

static double bicubic76(const view_t* v, const c8mi_native_bicubic* s, double e, double t) {
  if (s->polynomial_coefficient_offset != C8MI_NO_POLYNOMIAL_COEFFICIENTS &&
      v->bicubic_polynomial_coefficients != NULL)
    return bicubic_precomputed(v, s, e, t);
  return bicubic_sum(v, s, e, t, 0);
}

/* Six-point finite-difference loss derivative (:301-319). */
static double bicubic_loss_derivative76(const view_t* v, const c8mi_native_bicubic* s, double e,
                                        double t) {
  const double epsilon = 2.220446049250313080847263336181640625e-16;
  double h = pow(epsilon / 168., 1. / 7.);
  double coordinate = axis_transform(&s->loss_axis, t);
  double representable = coordinate + h;
  h = representable - coordinate;
#define AT(c) bicubic76(v, s, e, axis_back_transform(&s->loss_axis, (c)))
  double y1 = AT(coordinate + h) - AT(coordinate - h);
  double y2 = AT(coordinate - 2. * h) - AT(coordinate + 2. * h);
  double y3 = AT(coordinate + 3. * h) - AT(coordinate - 3. * h);
#undef AT
  double cd = (y3 + 9. * y2 + 45. * y1) / (60. * h);
  return cd * axis_derivative(&s->loss_axis, t);
}

/* evaluate_cubic (:321-338) and its derivative (:340-364). */
static double cubic(const view_t* v, const c8mi_native_cubic* s, double x) {
  double coordinate = axis_transform(&s->axis, x);
  uint32_t node = interval_node(coordinate, s->coefficient_count);
  double fraction = coordinate - node;
  uint64_t off = s->coefficient_offset + node;
  double b[4];
  hermite(fraction, b);
  double value = b[0] * v->cubic_values[off] + b[1] * v->cubic_values[off + 1] +
                 b[2] * v->cubic_node_derivatives[off] + b[3] * v->cubic_node_derivatives[off + 1];
  if (s->has_value_transform) value = axis_back_transform(&s->value_transform, value);
  return value;
}
static double cubic_derivative76(const view_t* v, const c8mi_native_cubic* s, double x) {
  double coordinate = axis_transform(&s->axis, x);
  uint32_t node = interval_node(coordinate, s->coefficient_count);
  double fraction = coordinate - node;
  uint64_t off = s->coefficient_offset + node;
  double b[4], d[4];
  hermite(fraction, b);
  hermite_d(fraction, d);
  double raw = b[0] * v->cubic_values[off] + b[1] * v->cubic_values[off + 1] +
               b[2] * v->cubic_node_derivatives[off] + b[3] * v->cubic_node_derivatives[off + 1];
  double der = d[0] * v->cubic_values[off] + d[1] * v->cubic_values[off + 1] +
               d[2] * v->cubic_node_derivatives[off] + d[3] * v->cubic_node_derivatives[off + 1];
  if (s->has_value_transform) der *= axis_back_derivative(&s->value_transform, raw);
  return der * axis_derivative(&s->axis, x);
}

static int sgn(double x) { return (0. < x ? 1 : 0) - (x < 0. ? 1 : 0); }

/* Utility-inversion residual context for proposal76_newton (:372-449). */
typedef struct {
  const view_t* v;
  const c8mi_native_cubic* s;
  double target;
} cubic_root_t;
static double cubic_residual(const cubic_root_t* c, double e) { return cubic(c->v, c->s, e) - c->target; }
static double cubic_slope(const cubic_root_t* c, double e) { return cubic_derivative76(c->v, c->s, e); }

typedef struct {
  double value;
  uint32_t iterations;
  int runtime_error;
} newton_t;

/* proposal76_newton: Boost 1.79 newton_raphson_iterate(digits = 20). */
static newton_t newton76(const cubic_root_t* ctx, double initial_guess, double lower,
                         double upper) {
  newton_t out;
  double residual = 0., last_residual = 0., result = initial_guess, guess = initial_guess;
  double delta = 1.7976931348623157e308, delta1 = delta, delta2 = delta;
  double maximum_residual = 0., minimum_residual = 0.;
  const double factor = 1.9073486328125e-6;
  uint32_t it = 0;
  while (it < 256) {
    last_residual = residual;
    delta2 = delta1;
    delta1 = delta;
    residual = cubic_residual(ctx, result);
    double derivative = cubic_slope(ctx, result);
    ++it;
    if (!isfinite(residual) || !isfinite(derivative)) {
      out.value = result; out.iterations = it; out.runtime_error = 1;
      return out;
    }
    if (residual == 0.) break;
    if (derivative == 0.) {
      if (last_residual == 0.) {
        guess = result == lower ? upper : lower;
        last_residual = cubic_residual(ctx, guess);
        delta = guess - result;
      }
      if (sgn(last_residual) * sgn(residual) < 0)
        delta = delta < 0. ? (result - lower) / 2. : (result - upper) / 2.;
      else
        delta = delta < 0. ? (result - upper) / 2. : (result - lower) / 2.;
    } else {
      delta = residual / derivative;
    }
    double abs_delta2 = delta2 < 0. ? -delta2 : delta2;
    if ((delta * 2. < 0. ? -(delta * 2.) : delta * 2.) > abs_delta2) {
      double shift = delta > 0. ? (result - lower) / 2. : (result - upper) / 2.;
      double abs_shift = shift < 0. ? -shift : shift;
      double abs_result = result < 0. ? -result : result;
      if (result != 0. && abs_shift > abs_result)
        delta = (delta > 0. ? 1. : -1.) * abs_result;
      else
        delta = shift;
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
    if (maximum_residual * minimum_residual > 0.) {
      out.value = result; out.iterations = it; out.runtime_error = 1;
      return out;
    }
    double abs_result = result < 0. ? -result : result;
    double abs_step = delta < 0. ? -delta : delta;
    if (!(abs_result * factor < abs_step)) break;
  }
  out.value = result;
  out.iterations = it;
  out.runtime_error = it == 256;
  return out;
}

/* valid_view (:451-468) */
static int valid_view(const view_t* v) {
  return v->dndx_column_count > 0 && v->dndx_columns != NULL && v->total_rate_column_count > 0 &&
         v->total_rate_columns != NULL && v->selection_column_indices != NULL &&
         (v->dedx_column_count == 0 ||
          (v->dedx_columns != NULL && v->dedx_accumulation_indices != NULL)) &&
         v->bicubic_values != NULL && v->bicubic_derivative_energy != NULL &&
         v->bicubic_derivative_loss != NULL && v->bicubic_mixed_derivative != NULL &&
         v->bicubic_coefficient_count > 0;
}

/* find_dndx: lower bound in canonical (pid, process, component) order (:470-500) */
static const c8mi_native_dndx_column* find_dndx(const view_t* v, int32_t pid, int32_t process,
                                                uint64_t component) {
  uint32_t lo = 0, hi = v->dndx_column_count;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    const c8mi_native_dndx_column* c = &v->dndx_columns[mid];
    int before = c->pdg_id < pid ||
                 (c->pdg_id == pid && (c->process_id < process ||
                                       (c->process_id == process && c->component_hash < component)));
    if (before) lo = mid + 1; else hi = mid;
  }
  if (lo < v->dndx_column_count) {
    const c8mi_native_dndx_column* c = &v->dndx_columns[lo];
    if (c->pdg_id == pid && c->process_id == process && c->component_hash == component) return c;
  }
  return NULL;
}

static void dndx_range(const view_t* v, int32_t pid, uint32_t* begin, uint32_t* end) {
  uint32_t lo = 0, hi = v->dndx_column_count;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (v->dndx_columns[mid].pdg_id < pid) lo = mid + 1; else hi = mid;
  }
  *begin = lo;
  hi = v->dndx_column_count;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (v->dndx_columns[mid].pdg_id <= pid) lo = mid + 1; else hi = mid;
  }
  *end = lo;
}

static const c8mi_native_dndx_column* selection_dndx(const view_t* v, uint32_t position) {
  if (position >= v->dndx_column_count) return NULL;
  uint32_t index = v->selection_column_indices[position];
  if (index >= v->dndx_column_count) return NULL;
  return &v->dndx_columns[index];
}

static const c8mi_native_total_rate_column* find_total_rate(const view_t* v, int32_t pid) {
  uint32_t lo = 0, hi = v->total_rate_column_count;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (v->total_rate_columns[mid].pdg_id < pid) lo = mid + 1; else hi = mid;
  }
  if (lo < v->total_rate_column_count && v->total_rate_columns[lo].pdg_id == pid)
    return &v->total_rate_columns[lo];
  return NULL;
}

static const c8mi_native_utility_column* find_utility(const view_t* v, int32_t pid) {
  for (uint32_t i = 0; i < v->utility_column_count; ++i)
    if (v->utility_columns[i].pdg_id == pid) return &v->utility_columns[i];
  return NULL;
}

typedef struct { double minimum, maximum; int valid; } limits_t;

/* kinematic_limits (:571-668) */
static limits_t kinematic_limits(const view_t* v, const c8mi_native_dndx_column* c, double e) {
  limits_t none = {0., 0., 0};
  if (!(e > 0.)) return none;
  double mass = c->particle_mass_mev, minimum = 0., maximum = 0.;
  const c8mi_native_constants* k = &v->constants;
  switch (c->kinematic_model) {
    case C8MI_KIN_ONLY_STOCHASTIC:
    case C8MI_KIN_PHOTON_PAIR_FULL_LOSS: {
      limits_t one = {1., 1., 1};
      return one;
    }
    case C8MI_KIN_COMPTON:
      maximum = 1. - 1. / (1. + 2. * e / k->electron_mass_mev);
      break;
    case C8MI_KIN_BREMS_ELECTRON_SCREENING:
      if (!(mass > 0.)) return none;
      maximum = 1. - mass / e;
      if (maximum < 0.) maximum = 0.;
      break;
    case C8MI_KIN_BREMS_KKP:
      if (!(mass > 0.) || !(c->component_nuclear_charge > 0.)) return none;
      maximum = 1. - 0.75 * k->sqrt_e * (mass / e) * pow(c->component_nuclear_charge, 1. / 3.);
      if (maximum < 0.) maximum = 0.;
      break;
    case C8MI_KIN_ELECTRON_PAIR: {
      if (!(mass > 0.)) return none;
      double ratio = mass / e;
      minimum = 4. * k->electron_mass_mev / e;
      maximum = 1. - 0.75 * k->sqrt_e * ratio * pow(c->component_nuclear_charge, 1. / 3.);
      double second = 1. - 6. * ratio * ratio;
      if (second < maximum) maximum = second;
      if (maximum < minimum) maximum = minimum;
      break;
    }
    case C8MI_KIN_PHOTONUCLEAR_ALLM97: {
      double nucleon = c->component_average_nucleon_mass_mev, pion = k->charged_pion_mass_mev;
      if (!(nucleon > 0.) || !(pion > 0.) || !(mass > 0.)) return none;
      minimum = (pion + pion * pion / (2. * nucleon)) / e;
      maximum = 1.;
      if (mass < pion) {
        double ratio = mass / nucleon;
        maximum -= nucleon * (1. + ratio * ratio) / (2. * e);
      }
      if (maximum < minimum) maximum = minimum;
      break;
    }
    case C8MI_KIN_IONIZATION_MOLLER:
      if (!(mass > 0.)) return none;
      maximum = 0.5 * (1. - mass / e);
      if (maximum < 0.) maximum = 0.;
      break;
    case C8MI_KIN_IONIZATION_BHABHA:
      if (!(mass > 0.)) return none;
      maximum = 1. - mass / e;
      if (maximum < 0.) maximum = 0.;
      break;
    case C8MI_KIN_IONIZATION_BETHE_BLOCH_ROSSI: {
      if (!(mass > 0.) || !(c->medium_mean_excitation_energy_mev > 0.)) return none;
      double mass_ratio = k->electron_mass_mev / mass, gamma = e / mass;
      minimum = c->medium_mean_excitation_energy_mev / e;
      maximum = 2. * k->electron_mass_mev * (gamma * gamma - 1.) /
                ((1. + 2. * gamma * mass_ratio + mass_ratio * mass_ratio) * e);
      double kinetic = 1. - mass / e;
      if (kinetic < maximum) maximum = kinetic;
      if (maximum < minimum) maximum = minimum;
      break;
    }
    default:
      return none;
  }
  if (!isfinite(minimum) || !isfinite(maximum) || maximum < minimum) return none;
  double cut = minimum;
  if (c->absolute_cut_mev > 0. && c->relative_cut > 0.) {
    double requested = c->absolute_cut_mev / e;
    if (c->relative_cut < requested) requested = c->relative_cut;
    if (requested < minimum) requested = minimum;
    if (requested > maximum) requested = maximum;
    cut = requested;
  }
  limits_t out = {cut, maximum, 1};
  return out;
}

/* physical_loss (:670-702) */
static double physical_loss(const view_t* v, const c8mi_native_dndx_column* c, double e, double t) {
  limits_t l = kinematic_limits(v, c, e);
  if (!l.valid) return NAN;
  double lower = l.minimum, upper = l.maximum;
  if (t <= 0.) return lower;
  if (t >= 1.) return upper;
  if (lower == upper) return lower;
  switch (c->loss_transform) {
    case C8MI_LOSS_RELATIVE_LOG:
      if (!(lower > 0.)) return NAN;
      return lower * exp(t * log(upper / lower));
    case C8MI_LOSS_RELATIVE_LOG_POW15:
      if (!(lower > 0.)) return NAN;
      return lower * exp(pow(t, 1.5) * log(upper / lower));
    case C8MI_LOSS_ONE_MINUS_LOG: {
      double xi = log((1. - lower) / (1. - upper));
      return 1. - (1. - upper) * exp((1. - t) * xi);
    }
  }
  return NAN;
}

/* transformed_loss (:704-735) */
static double transformed_loss(const view_t* v, const c8mi_native_dndx_column* c, double e,
                               double physical) {
  limits_t l = kinematic_limits(v, c, e);
  if (!l.valid || physical < l.minimum || physical > l.maximum) return NAN;
  double lower = l.minimum, upper = l.maximum;
  if (physical <= lower) return 0.;
  if (physical >= upper) return 1.;
  if (lower == upper) return 0.;
  switch (c->loss_transform) {
    case C8MI_LOSS_RELATIVE_LOG:
      if (!(lower > 0.)) return NAN;
      return log(physical / lower) / log(upper / lower);
    case C8MI_LOSS_RELATIVE_LOG_POW15: {
      if (!(lower > 0.)) return NAN;
      double ratio = log(physical / lower) / log(upper / lower);
      return pow(ratio, 2. / 3.);
    }
    case C8MI_LOSS_ONE_MINUS_LOG: {
      double xi = log((1. - lower) / (1. - upper));
      return 1. - log((1. - physical) / (1. - upper)) / xi;
    }
  }
  return NAN;
}

static int rate_below_threshold(const c8mi_native_dndx_column* c, double e) {
  return c->rate_model == C8MI_RATE_BICUBIC &&
         (e < c->lower_energy_limit_mev || e < c->spline.energy_axis.low);
}

/* query_proposal_native_cumulative_rate (:744-771) */
static result_t cumulative_rate(const view_t* v, const c8mi_native_dndx_column* c, double e,
                                double t) {
  if (!valid_view(v)) return res(C8MI_QUERY_INVALID_VIEW, 0, 0, 0.);
  if (!isfinite(e) || !isfinite(t)) return res(C8MI_QUERY_NON_FINITE_INPUT, 0, 0, 0.);
  if (e < c->lower_energy_limit_mev || e < c->spline.energy_axis.low)
    return res(C8MI_QUERY_SUCCESS, 0, 0, 0.);
  if (e > c->spline.energy_axis.high) return res(C8MI_QUERY_ENERGY_OUT_OF_RANGE, 0, 0, 0.);
  if (t < c->spline.loss_axis.low || t > c->spline.loss_axis.high)
    return res(C8MI_QUERY_QUANTILE_OUT_OF_RANGE, 0, 0, 0.);
  double value = bicubic76(v, &c->spline, e, t) / c->component_weight;
  if (value < 0.) value = 0.;
  if (!isfinite(value)) return res(C8MI_QUERY_INVALID_VIEW, 0, 0, 0.);
  return res(C8MI_QUERY_SUCCESS, 0, 0, value);
}

/* query_proposal_native_rate_for_column (:773-894) */
static result_t rate_for_column(const view_t* v, const c8mi_native_dndx_column* c, double e) {
  if (!valid_view(v)) return res(C8MI_QUERY_INVALID_VIEW, 0, 0, 0.);
  if (rate_below_threshold(c, e)) return res(C8MI_QUERY_SUCCESS, 0, 0, 0.);
  if (c->rate_model == C8MI_RATE_ANNIHILATION_HEITLER) {
    if (!isfinite(e)) return res(C8MI_QUERY_NON_FINITE_INPUT, 0, 0, 0.);
    if (e <= c->particle_mass_mev) return res(C8MI_QUERY_SUCCESS, 0, 0, 0.);
    double gamma = e / c->particle_mass_mev;
    double root = pow(gamma * gamma - 1., .5);
    double aux = (gamma * gamma + 4. * gamma + 1.) / (gamma * gamma - 1.) * log(gamma + root) -
                 (gamma + 3.) / root;
    return res(C8MI_QUERY_SUCCESS, 0, 0, aux * c->direct_rate_factor / (gamma + 1.));
  }
  if (c->rate_model == C8MI_RATE_PHOTOEFFECT_SAUTER) {
    if (!isfinite(e)) return res(C8MI_QUERY_NON_FINITE_INPUT, 0, 0, 0.);
    if (!(e > c->lower_energy_limit_mev)) return res(C8MI_QUERY_SUCCESS, 0, 0, 0.);
    double alpha = v->constants.fine_structure_constant, me = v->constants.electron_mass_mev;
    double pi = v->constants.pi, re = v->constants.classical_electron_radius_cm;
    double z = c->component_nuclear_charge;
    double ionization = z * z * alpha * alpha * me / 2.;
    double gamma = 1. + (e - ionization) / me;
    double gamma_beta = sqrt(gamma * gamma - 1.);
    double sigma_t = 8. * pi * re * re / 3.;
    double sigma = 1.5 * sigma_t * pow(z, 5.) * pow(alpha, 4.) * pow(me / e, 5.) *
                   pow(gamma_beta, 3.) *
                   (4. / 3. + gamma * (gamma - 2.) / (gamma + 1.) *
                                  (1. - log(gamma + gamma_beta) / (gamma * gamma_beta)));
    double zab = z * alpha * gamma / gamma_beta;
    double zpab = pi * zab;
    double correction;
    if (zpab < 10.)
      correction = (1. + zab * zab) * zpab / sinh(zpab) * exp(zab * (pi - 4. * atan(1. / zab)));
    else
      correction = (1. + zab * zab) * 2. * zpab * exp(-4. * zab * atan(1. / zab));
    double log_z = log(z);
    double shell = 1. + 0.01481 * log_z * log_z - 0.000788 * log_z * log_z * log_z;
    return res(C8MI_QUERY_SUCCESS, 0, 0, sigma * correction * shell * c->direct_rate_factor);
  }
  if (c->rate_model == C8MI_RATE_PHOTOPRODUCTION_HECK) {
    if (!isfinite(e)) return res(C8MI_QUERY_NON_FINITE_INPUT, 0, 0, 0.);
    if (e < c->lower_energy_limit_mev) return res(C8MI_QUERY_SUCCESS, 0, 0, 0.);
    double eg = e * 1.e-3, mn = c->component_average_nucleon_mass_mev * 1.e-3;
    double s = mn * mn + 2. * mn * eg;
    const double s0 = 1.0761 * 1.0761;
    double cross;
    if (sqrt(s) < 19.39)
      cross = (73.7 * pow(s, 0.073) + 191.7 * pow(s, -0.602)) * sqrt(1. - s0 / s);
    else
      cross = 59.3 * pow(s, 0.093) + 120.2 * pow(s, -0.358);
    static const double rm[3] = {1.231, 1.515, 1.680}, rw[3] = {0.11, 0.11, 0.125};
    static const double rs[3] = {31.125, 25.567, 17.508}, rwin[3] = {0.17, 0.38, 0.38};
    for (int i = 0; i < 3; ++i) {
      double m2 = rm[i] * rm[i], w2 = rw[i] * rw[i];
      double bw = s / (eg * eg) * rs[i] * w2 * s / ((s - m2) * (s - m2) + w2 * s);
      double th = 0.;
      if (eg >= 0.152 + rwin[i]) th = 1.;
      else if (eg >= 0.152) th = (eg - 0.152) / rwin[i];
      cross += bw * th;
    }
    return res(C8MI_QUERY_SUCCESS, 0, 0,
               cross * pow(c->component_atomic_mass, 0.91) * c->direct_rate_factor);
  }
  if (c->rate_model == C8MI_RATE_BICUBIC && (c->pdg_id == 13 || c->pdg_id == -13)) {
    if (!isfinite(e)) return res(C8MI_QUERY_NON_FINITE_INPUT, 0, 0, 0.);
    if (e > c->spline.energy_axis.high) return res(C8MI_QUERY_ENERGY_OUT_OF_RANGE, 0, 0, 0.);
    double value = bicubic_precomputed(v, &c->spline, e, 1.) / c->component_weight;
    if (value < 0.) value = 0.;
    if (!isfinite(value)) return res(C8MI_QUERY_INVALID_VIEW, 0, 0, 0.);
    return res(C8MI_QUERY_SUCCESS, 0, 0, value);
  }
  return cumulative_rate(v, c, e, 1.);
}

static result_t query_rate(const view_t* v, int32_t pid, int32_t process, uint64_t component,
                           double e) {
  if (!valid_view(v)) return res(C8MI_QUERY_INVALID_VIEW, 0, 0, 0.);
  const c8mi_native_dndx_column* c = find_dndx(v, pid, process, component);
  if (!c) return res(C8MI_QUERY_COLUMN_NOT_FOUND, 0, 0, 0.);
  if (rate_below_threshold(c, e)) return res(C8MI_QUERY_SUCCESS, 0, 0, 0.);
  return rate_for_column(v, c, e);
}

/* query_proposal_native_total_rate (:909-931) */
static result_t total_rate(const view_t* v, int32_t pid, double e) {
  if (!valid_view(v)) return res(C8MI_QUERY_INVALID_VIEW, 0, 0, 0.);
  if (!isfinite(e)) return res(C8MI_QUERY_NON_FINITE_INPUT, 0, 0, 0.);
  const c8mi_native_total_rate_column* c = find_total_rate(v, pid);
  if (!c) return res(C8MI_QUERY_PARTICLE_NOT_FOUND, 0, 0, 0.);
  if (e < c->lower_energy_limit_mev) return res(C8MI_QUERY_SUCCESS, 0, 0, 0.);
  if (e < c->spline.axis.low || e > c->spline.axis.high)
    return res(C8MI_QUERY_ENERGY_OUT_OF_RANGE, 0, 0, 0.);
  double value = cubic(v, &c->spline, e);
  if (!isfinite(value)) return res(C8MI_QUERY_INVALID_VIEW, 0, 0, 0.);
  return res(C8MI_QUERY_SUCCESS, 0, 0, value < 0. ? 0. : value);
}

/* query_proposal_native_selection_total_rate (:937-956) */
static result_t selection_total_rate(const view_t* v, int32_t pid, double e) {
  if (!valid_view(v)) return res(C8MI_QUERY_INVALID_VIEW, 0, 0, 0.);
  double total = 0.;
  uint32_t begin, end;
  dndx_range(v, pid, &begin, &end);
  int found = begin != end;
  for (uint32_t i = begin; i < end; ++i) {
    const c8mi_native_dndx_column* c = selection_dndx(v, i);
    if (!c || c->pdg_id != pid) return res(C8MI_QUERY_INVALID_VIEW, 0, 0, 0.);
    if (rate_below_threshold(c, e)) continue;
    result_t r = rate_for_column(v, c, e);
    if (r.status != C8MI_QUERY_SUCCESS) return r;
    total += r.value;
  }
  if (!found) return res(C8MI_QUERY_PARTICLE_NOT_FOUND, 0, 0, 0.);
  return res(C8MI_QUERY_SUCCESS, 0, 0, total);
}

/* query_proposal_native_dedx (:958-986) */
static result_t dedx(const view_t* v, int32_t pid, double e) {
  if (!isfinite(e)) return res(C8MI_QUERY_NON_FINITE_INPUT, 0, 0, 0.);
  double total = 0.;
  int found = 0;
  for (uint32_t p = 0; p < v->dedx_column_count; ++p) {
    uint32_t index = v->dedx_accumulation_indices[p];
    if (index >= v->dedx_column_count) return res(C8MI_QUERY_INVALID_VIEW, 0, 0, 0.);
    const c8mi_native_dedx_column* c = &v->dedx_columns[index];
    if (c->pdg_id != pid) continue;
    found = 1;
    if (e < c->lower_energy_limit_mev) continue;
    if (e < c->spline.axis.low || e > c->spline.axis.high)
      return res(C8MI_QUERY_ENERGY_OUT_OF_RANGE, 0, 0, 0.);
    total += cubic(v, &c->spline, e) / c->component_weight;
  }
  if (!found) return res(C8MI_QUERY_PARTICLE_NOT_FOUND, 0, 0, 0.);
  return res(C8MI_QUERY_SUCCESS, 0, 0, total);
}

/* query_proposal_native_loss_fraction (:988-1281) */
typedef struct {
  const view_t* v;
  const c8mi_native_bicubic* s;
  double e, target;
} loss_root_t;
static double loss_residual(const loss_root_t* c, double t) {
  return bicubic76(c->v, c->s, c->e, t) - c->target;
}

static result_t loss_fraction(const view_t* v, int32_t pid, int32_t process, uint64_t component,
                              double e, double q, int photon_split) {
  if (!valid_view(v)) return res(C8MI_QUERY_INVALID_VIEW, 0, 0, 0.);
  if (!isfinite(e) || !isfinite(q)) return res(C8MI_QUERY_NON_FINITE_INPUT, 0, 0, 0.);
  if (q < 0. || q > 1.) return res(C8MI_QUERY_QUANTILE_OUT_OF_RANGE, 0, 0, 0.);
  const c8mi_native_dndx_column* c = find_dndx(v, pid, process, component);
  if (!c) return res(C8MI_QUERY_COLUMN_NOT_FOUND, 0, 0, 0.);
  int muon_split = photon_split && process == 1000000015 &&
                   c->kinematic_model == C8MI_KIN_ONLY_STOCHASTIC;
  if (photon_split &&
      (pid != 22 || (!muon_split && c->kinematic_model != C8MI_KIN_PHOTON_PAIR_FULL_LOSS) ||
       c->loss_transform != C8MI_LOSS_RELATIVE_LOG))
    return res(C8MI_QUERY_UNSUPPORTED_KINEMATICS, 0, 0, 0.);
  if ((c->kinematic_model == C8MI_KIN_ONLY_STOCHASTIC && !muon_split) ||
      (c->kinematic_model == C8MI_KIN_PHOTON_PAIR_FULL_LOSS && !photon_split))
    return res(C8MI_QUERY_SUCCESS, 0, 0, 1.);
  if (c->rate_model != C8MI_RATE_BICUBIC) return res(C8MI_QUERY_UNSUPPORTED_KINEMATICS, 0, 0, 0.);
  double xmin = v->constants.electron_mass_mev / e;
  double xmax = 1. - xmin;
  if (muon_split) {
    double threshold = 2. * v->constants.muon_mass_mev * v->constants.sqrt_e;
    if (!(v->constants.muon_mass_mev > 0.) || !(e > threshold))
      return res(C8MI_QUERY_ENERGY_OUT_OF_RANGE, 0, 0, 0.);
    double width = sqrt(1. - 2. * v->constants.sqrt_e * v->constants.muon_mass_mev / e);
    xmin = .5 * (1. - width);
    xmax = .5 * (1. + width);
  }
  if (photon_split &&
      (xmin >= .5 || e < c->spline.energy_axis.low || e > c->spline.energy_axis.high))
    return res(C8MI_QUERY_ENERGY_OUT_OF_RANGE, 0, 0, 0.);
  limits_t lim;
  if (photon_split) {
    lim.minimum = xmin; lim.maximum = xmax; lim.valid = 1;
  } else {
    lim = kinematic_limits(v, c, e);
  }
  if (!lim.valid) return res(C8MI_QUERY_UNSUPPORTED_KINEMATICS, 0, 0, 0.);
  if (lim.minimum == lim.maximum) return res(C8MI_QUERY_SUCCESS, 0, 0, lim.minimum);
  result_t total;
  if (!photon_split) {
    total = cumulative_rate(v, c, e, 1.);
  } else {
    double raw = bicubic76(v, &c->spline, e, 1.);
    if (!isfinite(raw)) total = res(C8MI_QUERY_INVALID_VIEW, 0, 0, 0.);
    else total = res(C8MI_QUERY_SUCCESS, 0, 0, raw < 0. ? 0. : raw);
  }
  if (total.status != C8MI_QUERY_SUCCESS) return total;
  double target = photon_split ? q * total.value : (q * total.value) * c->component_weight;
  if (q == 0.) return res(C8MI_QUERY_SUCCESS, 0, 0, lim.minimum);
  if (q == 1.) return res(C8MI_QUERY_SUCCESS, 0, 0, lim.maximum);
  loss_root_t ctx = {v, &c->spline, e, target};

  double lower = c->spline.loss_axis.low, upper = c->spline.loss_axis.high;
  double lower_residual = loss_residual(&ctx, lower);
  double upper_residual = loss_residual(&ctx, upper);
  uint32_t bisection = 0;
  int need_bisection = 0, exact = 0;
  double current = 0.;
  if (!isfinite(lower_residual) || !isfinite(upper_residual)) {
    return res(C8MI_QUERY_ROOT_NOT_CONVERGED, 0, 0, 0.);
  } else if (lower_residual == 0.) {
    current = lower; exact = 1;
  } else if (upper_residual == 0.) {
    current = upper; exact = 1;
  } else if (lower_residual * upper_residual >= 0.) {
    need_bisection = 1;
  } else {
    double coarse = (upper - lower) * 1.e-2;
    while (!((upper - lower) < coarse)) {
      double middle = (lower + upper) / 2.;
      double middle_residual = loss_residual(&ctx, middle);
      ++bisection;
      if (!isfinite(middle_residual)) { need_bisection = 1; break; }
      if (middle == upper || middle == lower) break;
      if (middle_residual == 0.) { lower = upper = middle; exact = 1; break; }
      if (sgn(middle_residual) * sgn(lower_residual) < 0) {
        upper = middle; upper_residual = middle_residual;
      } else {
        lower = middle; lower_residual = middle_residual;
      }
    }
    current = (lower + upper) / 2.;
  }

  double delta = 1.7976931348623157e308, delta1 = delta, delta2 = delta;
  double previous_residual = 0., maximum_residual = 0., minimum_residual = 0.;
  const double factor = 1.9073486328125e-6;
  uint32_t newton = 0;
  double guess = current;
  while (!exact && !need_bisection && newton < 256) {
    double last_residual = previous_residual;
    delta2 = delta1;
    delta1 = delta;
    double residual = loss_residual(&ctx, current);
    previous_residual = residual;
    ++newton;
    if (!isfinite(residual)) { need_bisection = 1; break; }
    if (residual == 0.) { exact = 1; break; }
    double derivative = bicubic_loss_derivative76(v, &c->spline, e, current);
    if (!isfinite(derivative)) {
      need_bisection = 1;
      break;
    } else if (derivative == 0.) {
      double mutable_last = last_residual;
      if (mutable_last == 0.) {
        guess = current == lower ? upper : lower;
        mutable_last = loss_residual(&ctx, guess);
        delta = guess - current;
      }
      if (sgn(mutable_last) * sgn(residual) < 0)
        delta = delta < 0. ? (current - lower) / 2. : (current - upper) / 2.;
      else
        delta = delta < 0. ? (current - upper) / 2. : (current - lower) / 2.;
    } else {
      delta = residual / derivative;
    }
    double abs_delta = delta < 0. ? -delta : delta;
    double abs_delta2 = delta2 < 0. ? -delta2 : delta2;
    if (2. * abs_delta > abs_delta2) {
      double shift = delta > 0. ? (current - lower) * .5 : (current - upper) * .5;
      double abs_shift = shift < 0. ? -shift : shift;
      double abs_current = current < 0. ? -current : current;
      if (current != 0. && abs_shift > abs_current)
        delta = (delta > 0. ? 1. : -1.) * abs_current;
      else
        delta = shift;
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
      upper = guess; maximum_residual = residual;
    } else {
      lower = guess; minimum_residual = residual;
    }
    if (maximum_residual * minimum_residual > 0.) { need_bisection = 1; break; }
    double abs_current = current < 0. ? -current : current;
    double abs_step = delta < 0. ? -delta : delta;
    if (!(abs_current * factor < abs_step)) break;
  }
  if (newton == 256) need_bisection = 1;

  if (need_bisection) {
    lower = c->spline.loss_axis.low;
    upper = c->spline.loss_axis.high;
    lower_residual = loss_residual(&ctx, lower);
    upper_residual = loss_residual(&ctx, upper);
    if (!isfinite(lower_residual) || !isfinite(upper_residual) ||
        lower_residual * upper_residual > 0.)
      return res(C8MI_QUERY_ROOT_NOT_CONVERGED, newton, bisection, 0.);
    for (uint32_t it = 0; it <= 100; ++it) {
      double middle = (lower + upper) / 2.;
      double middle_residual = loss_residual(&ctx, middle);
      double current_lower = loss_residual(&ctx, lower);
      if (!isfinite(middle_residual) || !isfinite(current_lower))
        return res(C8MI_QUERY_ROOT_NOT_CONVERGED, newton, bisection, 0.);
      if (sgn(middle_residual) == sgn(current_lower)) lower = middle; else upper = middle;
      ++bisection;
      if ((upper - lower) < 1.e-6) break;
    }
    current = (lower + upper) / 2.;
  }
  double physical = photon_split
                        ? lim.minimum * exp(current * log(lim.maximum / lim.minimum))
                        : physical_loss(v, c, e, current);
  if (!isfinite(physical)) return res(C8MI_QUERY_ROOT_NOT_CONVERGED, newton, bisection, 0.);
  return res(C8MI_QUERY_SUCCESS, newton, bisection, physical);
}

/* query_proposal_native_range (:1283-1301) */
static result_t range(const view_t* v, int32_t pid, double e) {
  const c8mi_native_utility_column* c = find_utility(v, pid);
  if (!c) return res(C8MI_QUERY_PARTICLE_NOT_FOUND, 0, 0, 0.);
  if (!isfinite(e)) return res(C8MI_QUERY_NON_FINITE_INPUT, 0, 0, 0.);
  if (e < c->lower_energy_limit_mev || e < c->spline.axis.low || e > c->spline.axis.high)
    return res(C8MI_QUERY_ENERGY_OUT_OF_RANGE, 0, 0, 0.);
  double upper = cubic(v, &c->spline, e);
  double lower = cubic(v, &c->spline, c->lower_energy_limit_mev);
  double value = c->reverse ? lower - upper : upper - lower;
  if (!isfinite(value) || value < 0.) return res(C8MI_QUERY_INVALID_VIEW, 0, 0, 0.);
  return res(C8MI_QUERY_SUCCESS, 0, 0, value);
}

/* PROPOSAL Bisection on the utility residual (:1350-1372). Returns 0 on failure. */
static int utility_bisection(const cubic_root_t* ctx, double accuracy, double* lower,
                             double* upper, uint32_t* iterations) {
  double lv = cubic_residual(ctx, *lower), uv = cubic_residual(ctx, *upper);
  if (!isfinite(lv) || !isfinite(uv) || lv * uv > 0.) return 0;
  for (uint32_t it = 0; it <= 100; ++it) {
    double center = (*lower + *upper) / 2.;
    double cv = cubic_residual(ctx, center);
    double clv = cubic_residual(ctx, *lower);
    if (!isfinite(cv) || !isfinite(clv)) return 0;
    if (sgn(cv) == sgn(clv)) *lower = center; else *upper = center;
    ++*iterations;
    if ((*upper - *lower) < accuracy) return 1;
  }
  return 1;
}

/* query_proposal_native_energy_after_continuous_loss (:1310-1395) */
static result_t energy_after_loss(const view_t* v, int32_t pid, double e0, double grammage) {
  const c8mi_native_utility_column* c = find_utility(v, pid);
  if (!c) return res(C8MI_QUERY_PARTICLE_NOT_FOUND, 0, 0, 0.);
  if (!isfinite(e0) || !isfinite(grammage) || grammage < 0.)
    return res(C8MI_QUERY_NON_FINITE_INPUT, 0, 0, 0.);
  if (e0 < c->lower_energy_limit_mev || e0 > c->spline.axis.high)
    return res(C8MI_QUERY_ENERGY_OUT_OF_RANGE, 0, 0, 0.);
  result_t maximum = range(v, pid, e0);
  if (maximum.status != C8MI_QUERY_SUCCESS) return maximum;
  if (grammage > maximum.value) return res(C8MI_QUERY_ENERGY_OUT_OF_RANGE, 0, 0, 0.);
  if (grammage == maximum.value) return res(C8MI_QUERY_SUCCESS, 0, 0, c->lower_energy_limit_mev);
  if (grammage == 0.) return res(C8MI_QUERY_SUCCESS, 0, 0, e0);
  double signed_grammage = c->reverse ? -grammage : grammage;
  double to_upper = cubic(v, &c->spline, e0);
  cubic_root_t ctx = {v, &c->spline, to_upper - signed_grammage};
  double lower = c->lower_energy_limit_mev, upper = e0;
  uint32_t bisection = 0;
  double coarse = (e0 - c->lower_energy_limit_mev) * 1.e-2;
  if (!utility_bisection(&ctx, coarse, &lower, &upper, &bisection))
    return res(C8MI_QUERY_ROOT_NOT_CONVERGED, 0, bisection, 0.);
  double initial = upper == e0 ? e0 : (lower + upper) / 2.;
  newton_t n = newton76(&ctx, initial, lower, upper);
  if (!n.runtime_error && isfinite(n.value))
    return res(C8MI_QUERY_SUCCESS, n.iterations, bisection, n.value);
  lower = c->lower_energy_limit_mev;
  upper = e0;
  if (!utility_bisection(&ctx, 1.e-6, &lower, &upper, &bisection))
    return res(C8MI_QUERY_ROOT_NOT_CONVERGED, n.iterations, bisection, 0.);
  return res(C8MI_QUERY_SUCCESS, n.iterations, bisection, lower);
}

/* query_proposal_native_energy (:1397-1428) */
static result_t energy_of_range(const view_t* v, int32_t pid, double r) {
  const c8mi_native_utility_column* c = find_utility(v, pid);
  if (!c) return res(C8MI_QUERY_PARTICLE_NOT_FOUND, 0, 0, 0.);
  if (!isfinite(r) || r < 0.) return res(C8MI_QUERY_NON_FINITE_INPUT, 0, 0, 0.);
  result_t maximum = range(v, pid, c->spline.axis.high);
  if (maximum.status != C8MI_QUERY_SUCCESS) return maximum;
  if (r > maximum.value) return res(C8MI_QUERY_ENERGY_OUT_OF_RANGE, 0, 0, 0.);
  double lower = c->lower_energy_limit_mev, upper = c->spline.axis.high;
  uint32_t it = 0;
  for (; it < 100; ++it) {
    double lc = axis_transform(&c->spline.axis, lower);
    double uc = axis_transform(&c->spline.axis, upper);
    double middle = axis_back_transform(&c->spline.axis, 0.5 * (lc + uc));
    result_t value = range(v, pid, middle);
    if (value.status != C8MI_QUERY_SUCCESS) return value;
    if (value.value < r) lower = middle; else upper = middle;
    if ((upper - lower) <= 1.e-11 * upper) break;
  }
  return res(C8MI_QUERY_SUCCESS, 0, it, 0.5 * (lower + upper));
}

void oracle_native_queries(const c8mi_native_table_set* table, const c8mi_native_query* queries,
                           uint64_t count, c8mi_native_result* out) {
  for (uint64_t i = 0; i < count; ++i) {
    const c8mi_native_query* q = &queries[i];
    switch (q->kind) {
      case C8MI_NATIVE_TOTAL_RATE: out[i] = total_rate(table, q->pdg_id, q->energy_mev); break;
      case C8MI_NATIVE_SELECTION_TOTAL_RATE:
        out[i] = selection_total_rate(table, q->pdg_id, q->energy_mev);
        break;
      case C8MI_NATIVE_RATE:
        out[i] = query_rate(table, q->pdg_id, q->process_id, q->component_hash, q->energy_mev);
        break;
      case C8MI_NATIVE_DEDX: out[i] = dedx(table, q->pdg_id, q->energy_mev); break;
      case C8MI_NATIVE_LOSS_FRACTION:
      case C8MI_NATIVE_PHOTON_SPLIT:
        out[i] = loss_fraction(table, q->pdg_id, q->process_id, q->component_hash, q->energy_mev,
                               q->argument, q->kind == C8MI_NATIVE_PHOTON_SPLIT);
        break;
      case C8MI_NATIVE_RANGE: out[i] = range(table, q->pdg_id, q->energy_mev); break;
      case C8MI_NATIVE_ENERGY_AFTER_LOSS:
        out[i] = energy_after_loss(table, q->pdg_id, q->energy_mev, q->argument);
        break;
      case C8MI_NATIVE_ENERGY: out[i] = energy_of_range(table, q->pdg_id, q->argument); break;
      case C8MI_NATIVE_CUMULATIVE_RATE: {
        const c8mi_native_dndx_column* c =
            find_dndx(table, q->pdg_id, q->process_id, q->component_hash);
        out[i] = c ? cumulative_rate(table, c, q->energy_mev, q->argument)
                   : res(C8MI_QUERY_COLUMN_NOT_FOUND, 0, 0, 0.);
        break;
      }
      default: out[i] = res(C8MI_QUERY_INVALID_VIEW, 0, 0, 0.);
    }
  }
}

/* Exposed helpers for the tests. */
double oracle_native_physical_loss(const c8mi_native_table_set* v, uint32_t column, double e,
                                   double t) {
  return physical_loss(v, &v->dndx_columns[column], e, t);
}
double oracle_native_transformed_loss(const c8mi_native_table_set* v, uint32_t column, double e,
                                      double x) {
  return transformed_loss(v, &v->dndx_columns[column], e, x);
}

/* ---------------- entry points for the EM step restatement (em_oracle.c) ---------------- */
#include "native_internal.h"

result_t oracle_nq_total_rate(const view_t* v, int32_t pid, double e) {
  return total_rate(v, pid, e);
}
result_t oracle_nq_selection_total_rate(const view_t* v, int32_t pid, double e) {
  return selection_total_rate(v, pid, e);
}
result_t oracle_nq_rate_for_column(const view_t* v, const c8mi_native_dndx_column* c, double e) {
  return rate_for_column(v, c, e);
}
int oracle_nq_rate_below_threshold(const c8mi_native_dndx_column* c, double e) {
  return rate_below_threshold(c, e);
}
result_t oracle_nq_loss_fraction(const view_t* v, int32_t pid, int32_t process, uint64_t component,
                                 double e, double q, int photon_split) {
  return loss_fraction(v, pid, process, component, e, q, photon_split);
}
result_t oracle_nq_range(const view_t* v, int32_t pid, double e) { return range(v, pid, e); }
result_t oracle_nq_energy_after_loss(const view_t* v, int32_t pid, double e0, double grammage) {
  return energy_after_loss(v, pid, e0, grammage);
}
const c8mi_native_utility_column* oracle_nq_find_utility(const view_t* v, int32_t pid) {
  return find_utility(v, pid);
}
void oracle_nq_dndx_range(const view_t* v, int32_t pid, uint32_t* begin, uint32_t* end) {
  dndx_range(v, pid, begin, end);
}
const c8mi_native_dndx_column* oracle_nq_selection_dndx(const view_t* v, uint32_t position) {
  return selection_dndx(v, position);
}
