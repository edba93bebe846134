// Molière multiple scattering of one e+- step on the device: the projected-angle
// distribution with the analytical correction functions or the interpolation
// table, two Newton-solved projected angles, and the deflection of the step's
// end direction.
//
// Reference: corsika/detail/kokkos/em/models/MoliereScattering.inl:118-688
// (evaluate_interpolation, improve_moliere_initial_guess, normal_ppf,
// solve_moliere_b, f1_m / f2_m and their integrals with and without the
// interpolation view, build_distribution, evaluate_distribution,
// sample_one_dimensional, apply_moliere_direction and its step-compatible form)
// and corsika/detail/kokkos/em/algorithms/MoliereStep.inl:11-101.
// The snapshot (c8mi_moliere_snapshot) sits in the EmContext and is read through
// a uniform pointer, so its coefficients are scalar loads shared by the wavefront.
#pragma once

#include <hip/hip_runtime.h>

#include "em/hd.hpp"

#include <cstdint>

#include "c8mi/c8mi_em.h"

namespace c8mi::em::moliere {

enum Status : int { kSuccess = 0, kNoDeflection = 1, kInvalidSnapshot = 2, kInvalidInput = 3,
                    kNonConvergent = 4, kNonFinite = 5 };

inline constexpr double kPi = 3.141592653589793238462643383279502884;
inline constexpr double kSqrt2 = 1.414213562373095048801688724209698079;
inline constexpr double kMinimumB = 4.5;
inline constexpr double kMinimumBRhs = 2.9959226032237258;

// Per-step distribution state for a medium of Cap components (loops over the
// components are unrolled so the state stays in registers).
template <int Cap>
struct State {
  double inverse_scale[Cap];
  double inverse_b[Cap];
  double density_weight[Cap];
  double prefactor;
  double beta_squared;
};

// Acklam's rational approximation of the normal quantile (the Newton seed).
C8MI_HD double normal_ppf(double p) {
  constexpr double a0 = -3.969683028665376e+01, a1 = 2.209460984245205e+02, a2 = -2.759285104469687e+02,
                   a3 = 1.383577518672690e+02, a4 = -3.066479806614716e+01, a5 = 2.506628277459239e+00;
  constexpr double b0 = -5.447609879822406e+01, b1 = 1.615858368580409e+02, b2 = -1.556989798598866e+02,
                   b3 = 6.680131188771972e+01, b4 = -1.328068155288572e+01;
  constexpr double c0 = -7.784894002430293e-03, c1 = -3.223964580411365e-01, c2 = -2.400758277161838e+00,
                   c3 = -2.549732539343734e+00, c4 = 4.374664141464968e+00, c5 = 2.938163982698783e+00;
  constexpr double d0 = 7.784695709041462e-03, d1 = 3.224671290700398e-01, d2 = 2.445134137142996e+00,
                   d3 = 3.754408661907416e+00;
  constexpr double lo = 0.02425, hi = 1. - lo;
  if (p < lo) {
    double const q = sqrt(-2. * log(p));
    return (((((c0 * q + c1) * q + c2) * q + c3) * q + c4) * q + c5) / ((((d0 * q + d1) * q + d2) * q + d3) * q + 1.);
  }
  if (p <= hi) {
    double const q = p - 0.5, r = q * q;
    return (((((a0 * r + a1) * r + a2) * r + a3) * r + a4) * r + a5) * q /
           (((((b0 * r + b1) * r + b2) * r + b3) * r + b4) * r + 1.);
  }
  double const q = sqrt(-2. * log(1. - p));
  return -(((((c0 * q + c1) * q + c2) * q + c3) * q + c4) * q + c5) / ((((d0 * q + d1) * q + d2) * q + d3) * q + 1.);
}

// B - ln B = rhs on the B > 1 branch: fourth-order asymptotic seed, one Halley step.
C8MI_HD double solve_b(double rhs) {
  if (!isfinite(rhs) || rhs < kMinimumBRhs) return 0.;
  double const l = log(rhs), ir = 1. / rhs, ir2 = ir * ir, l2 = l * l, l3 = l2 * l;
  double b = rhs + l + l * ir + l * (2. - l) * 0.5 * ir2 + l * (6. - 9. * l + 2. * l2) / 6. * ir2 * ir +
             l * (12. - 36. * l + 22. * l2 - 3. * l3) / 12. * ir2 * ir2;
  if (!isfinite(b) || !(b > 1.)) return 0.;
  double const f = b - log(b) - rhs, ib = 1. / b, df = 1. - ib;
  double const den = 2. * df * df - f * ib * ib;
  if (!isfinite(den) || den == 0.) return 0.;
  b -= 2. * f * df / den;
  return isfinite(b) && b >= kMinimumB ? b : 0.;
}

// Integral series use the quotients c[i] / (2 i + 1), -0.5 c2_large[i] / i and
// 0.5 / i divided once on the host (`q`), the values each term would compute.
struct Quotients {
  double const* i1;
  double const* i2;
  double const* i2_large;
  double const* half_over;
};

// Interpolation view (MoliereInterpolationView): the four functions' cubics on
// x in [0, 20] and the initial-guess corrections; `polynomials` null for the
// analytical functions.
struct Table {
  c8mi_moliere_cubic const* polynomials;
  double const* guess;
};

inline constexpr double kTableUpper = 20.;
inline constexpr int kIntervals = C8MI_MOLIERE_INTERVALS;

// The snapshot was validated when the backend was created.
template <int Cap>
C8MI_HD int build(c8mi_moliere_snapshot const& m, double grammage, double e_mev, State<Cap>& st) {
  if (!isfinite(grammage) || grammage < 0. || !isfinite(e_mev) || !(e_mev > m.particle_mass_mev))
    return kInvalidInput;
  if (grammage == 0.) return kNoDeflection;
  double const mass = m.particle_mass_mev;
  double const p2 = (e_mev - mass) * (e_mev + mass);
  double const beta2 = 1. / (1. + mass * mass / p2);
  st.beta_squared = beta2;
  double bp2 = p2 / e_mev;
  bp2 *= bp2;
  double const chi_c2 = 4. * kPi * m.avogadro_per_mol * m.fine_structure_constant * m.fine_structure_constant *
                        m.hbar_mev_s * m.hbar_mev_s * m.speed_of_light_cm_per_s * m.speed_of_light_cm_per_s *
                        grammage / bp2 * m.z_squared_over_a_average;
  if (!isfinite(chi_c2) || !(chi_c2 > 0.)) return kNonFinite;
  double b_max = 0.;
#pragma unroll
  for (int i = 0; i < Cap; ++i) {
    if (i >= static_cast<int>(m.component_count)) {  // unused chains evaluate at x = 0
      st.inverse_scale[i] = 0.;
      st.inverse_b[i] = 0.;
      st.density_weight[i] = 0.;
      continue;
    }
    double const chi_a2 =
        m.components[i].chi0_squared_mev2 / p2 * (1.13 + m.components[i].coulomb_correction / beta2);
    double const b = solve_b(log(chi_c2 / chi_a2) + 1. - 2. * m.euler_mascheroni);
    if (!(b > 0.)) return kNoDeflection;
    double const scale = chi_c2 * b;
    if (!isfinite(scale) || !(scale > 0.)) return kNonFinite;
    st.inverse_scale[i] = 1. / scale;
    st.inverse_b[i] = 1. / b;
    st.density_weight[i] = m.components[i].weight_zz / sqrt(scale * kPi);
    if (!isfinite(st.inverse_scale[i]) || !(st.inverse_scale[i] > 0.) || !isfinite(st.inverse_b[i]) ||
        !(st.inverse_b[i] > 0.) || !isfinite(st.density_weight[i]) || !(st.density_weight[i] > 0.))
      return kNonFinite;
    if (i == static_cast<int>(m.maximum_weight_index)) b_max = b;
  }
  st.prefactor = sqrt(chi_c2 * b_max);
  return isfinite(st.prefactor) && st.prefactor > 0. ? kSuccess : kNonFinite;
}

// Large-x branches of the four functions (the power series cover x <= 12 for
// f1 and its integral, x <= 4.25^2 for f2 and its integral): rare tails, kept
// out of line so the step kernel's code stays small.
C8MI_HD_NOINLINE double f1_large(double x) {
  return 0.5 * sqrt(kPi) / (pow(x, 1.5) * pow(1. - 4.5 / x, 2. / 3.));
}
C8MI_HD_NOINLINE double integral_f1_large(c8mi_moliere_snapshot const& m, double x) {
  double sum = m.c1_integral_large[14];
  for (int i = 13; i >= 0; --i) sum = m.c1_integral_large[i] + sum / x;
  return sum;
}
// f2 and its integral between 4.25^2 and 6.5^2 (fitted), and the terms of
// their expansions above.
C8MI_HD double f2_fit(double x) {
  return 0.00013567765224589459194769192063035 * x - 0.0022635502525409950842771866774683 * sqrt(x) +
         0.0098037758070269476889935233998585;
}
C8MI_HD double integral_f2_fit(double x) {
  return -0.00026360133958801203364619158975302 * x + 0.0039965027465457608410459577896745 * sqrt(x) -
         0.016305842044996649714549974419242;
}
C8MI_HD double f2_term(c8mi_moliere_snapshot const& m, double x, int i) {
  return m.c2_large[i] * (0.5 * log(x) + m.s2_large[i]) * pow(x, -(i + 0.5));
}
// pow(x, -i) with an int exponent is pown here (measured cheaper than pow with
// a double exponent in the step kernel).
C8MI_HD double integral_f2_term(c8mi_moliere_snapshot const& m, Quotients const& q, double x,
                                int i) {
  return q.i2_large[i] * (q.half_over[i] + 0.5 * log(x) + m.s2_large[i]) * pow(x, -i);
}
C8MI_HD_NOINLINE double f2_large(c8mi_moliere_snapshot const& m, double x) {
  if (x <= 6.5 * 6.5) return f2_fit(x);
  double sum = 0.;
  for (int i = 2; i < 13; ++i) sum += f2_term(m, x, i);
  return sum;
}
C8MI_HD_NOINLINE double integral_f2_large(c8mi_moliere_snapshot const& m, Quotients const& q, double x) {
  if (x <= 6.5 * 6.5) return integral_f2_fit(x);
  double sum = 0.;
  for (int i = 2; i < 13; ++i) sum += integral_f2_term(m, q, x, i);
  return sum;
}

// Density and CDF of one projected angle (evaluate_distribution). The four
// power series of all N components run in one loop over the coefficients:
// 4N independent Horner chains share each coefficient load and fill the FP64
// pipeline, each chain with the same operations in the same order as alone.
template <int N>
C8MI_HD void evaluate(c8mi_moliere_snapshot const& m, Quotients const& q, State<N> const& st,
                      double theta, double& density, double& cumulative) {
  double const t2 = theta * theta;
  double x[N], s1[N], s2[N], si1[N], si2[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    x[k] = t2 * st.inverse_scale[k];
    s1[k] = m.c1[69];
    s2[k] = m.c2[69];
    si1[k] = q.i1[69];
    si2[k] = q.i2[69];
  }
  for (int i = 68; i >= 0; --i) {
    double const a = m.c1[i], b = m.c2[i], c = q.i1[i], d = q.i2[i];
#pragma unroll
    for (int k = 0; k < N; ++k) {
      s1[k] = s1[k] * x[k] + a;
      s2[k] = s2[k] * x[k] + b;
      si1[k] = si1[k] * x[k] + c;
      si2[k] = si2[k] * x[k] + d;
    }
  }
  double dens = 0., cdf = 0.;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    if (k >= static_cast<int>(m.component_count)) break;
    double const xk = x[k], ib = st.inverse_b[k], ib2 = ib * ib;
    double const root = sqrt(xk > 0. ? xk : 0.);
    double const f1v = xk > 12. ? f1_large(xk) : s1[k];
    double const i1v = xk > 12. ? integral_f1_large(m, xk) : si1[k] * root;
    double const f2v = xk > 4.25 * 4.25 ? f2_large(m, xk) : s2[k];
    double const i2v = xk > 4.25 * 4.25 ? integral_f2_large(m, q, xk) : si2[k] * root;
    dens += st.density_weight[k] * (exp(-xk) + f1v * ib + f2v * ib2);
    cdf += m.components[k].weight_zz * (0.5 * erf(root) + sqrt(1. / kPi) * (i1v * ib + i2v * ib2));
  }
  density = dens * m.inverse_weight_zz_sum;
  cdf *= m.inverse_weight_zz_sum;
  cumulative = theta < 0. ? -cdf : cdf;
}

// evaluate_distribution with an interpolation view: the cubic of x's interval
// for x <= 20 (the four functions share the interval), the large-x branches
// above (x > 20 is past both branch points of the series).
template <int N>
C8MI_HD void evaluate_table(c8mi_moliere_snapshot const& m, Quotients const& q, Table const& t,
                            State<N> const& st, double theta, double& density,
                            double& cumulative) {
  double const t2 = theta * theta;
  double dens = 0., cdf = 0.;
#pragma unroll 1
  for (int k = 0; k < N; ++k) {
    if (k >= static_cast<int>(m.component_count)) break;
    double const x = t2 * st.inverse_scale[k], ib = st.inverse_b[k], ib2 = ib * ib;
    double f1v, f2v, i1v, i2v;
    if (x <= kTableUpper) {
      // evaluate_interpolation (MoliereScattering.inl:118-134)
      double coordinate = (x - 0.) * (static_cast<double>(kIntervals) / (kTableUpper - 0.));
      if (coordinate < 0.) coordinate = 0.;
      if (coordinate > static_cast<double>(kIntervals)) coordinate = static_cast<double>(kIntervals);
      auto interval = static_cast<std::uint32_t>(coordinate);
      if (interval >= static_cast<std::uint32_t>(kIntervals)) interval = kIntervals - 1;
      double const u = coordinate - static_cast<double>(interval);
      auto cubic = [&](int f) {
        double const* c = t.polynomials[f * kIntervals + interval].c;
        return c[0] + u * (c[1] + u * (c[2] + u * c[3]));
      };
      f1v = cubic(0);
      f2v = cubic(1);
      i1v = cubic(2);
      i2v = cubic(3);
    } else {
      f1v = f1_large(x);
      i1v = integral_f1_large(m, x);
      f2v = f2_large(m, x);
      i2v = integral_f2_large(m, q, x);
    }
    double const root = sqrt(x > 0. ? x : 0.);
    dens += st.density_weight[k] * (exp(-x) + f1v * ib + f2v * ib2);
    cdf += m.components[k].weight_zz * (0.5 * erf(root) + sqrt(1. / kPi) * (i1v * ib + i2v * ib2));
  }
  density = dens * m.inverse_weight_zz_sum;
  cdf *= m.inverse_weight_zz_sum;
  cumulative = theta < 0. ? -cdf : cdf;
}

// improve_moliere_initial_guess (MoliereScattering.inl:136-216): the Gaussian
// coordinate plus the trilinear interpolation of the tabulated correction.
C8MI_HD double improve_guess(double const* g, double reference_b, double beta2, double gaussian) {
  constexpr double b_lo = 4.5, b_hi = 64., beta_lo = 0.25, beta_hi = 1., c_lo = 0., c_hi = 5.;
  constexpr std::uint32_t nb = C8MI_MOLIERE_GUESS_B_NODES, nt = C8MI_MOLIERE_GUESS_BETA2_NODES,
                          nc = C8MI_MOLIERE_GUESS_COORDINATE_NODES;
  constexpr double inv_b = static_cast<double>(nb - 1) / (b_hi - b_lo);
  constexpr double inv_beta = static_cast<double>(nt - 1) / (beta_hi - beta_lo);
  constexpr double inv_c = static_cast<double>(nc - 1) / (c_hi - c_lo);
  if (!isfinite(reference_b) || !isfinite(beta2) || !isfinite(gaussian)) return gaussian;
  double const mag = fabs(gaussian);
  if (reference_b < b_lo || reference_b > b_hi || beta2 < beta_lo || beta2 > beta_hi || mag < c_lo || mag > c_hi)
    return gaussian;
  double const bc = (reference_b - b_lo) * inv_b, cc = (mag - c_lo) * inv_c, tc = (beta2 - beta_lo) * inv_beta;
  auto bi = static_cast<std::uint32_t>(bc), ti = static_cast<std::uint32_t>(tc), ci = static_cast<std::uint32_t>(cc);
  if (bi + 1 >= nb) bi = nb - 2;
  if (ti + 1 >= nt) ti = nt - 2;
  if (ci + 1 >= nc) ci = nc - 2;
  double const bf = bc - static_cast<double>(bi), tf = tc - static_cast<double>(ti), cf = cc - static_cast<double>(ci);
  std::uint32_t const beta_stride = nt * nc;
  std::uint32_t const lb = bi * beta_stride, ub = lb + beta_stride;
  std::uint32_t const lt = ti * nc + ci, ut = lt + nc;
  double const ll = g[lb + lt] * (1. - cf) + g[lb + lt + 1] * cf;
  double const lu = g[lb + ut] * (1. - cf) + g[lb + ut + 1] * cf;
  double const ul = g[ub + lt] * (1. - cf) + g[ub + lt + 1] * cf;
  double const uu = g[ub + ut] * (1. - cf) + g[ub + ut + 1] * cf;
  double const lower = ll * (1. - tf) + lu * tf;
  double const upper = ul * (1. - tf) + uu * tf;
  double const corrected = mag + lower * (1. - bf) + upper * bf;
  if (!isfinite(corrected) || corrected < 0.) return gaussian;
  return gaussian < 0. ? -corrected : corrected;
}

// Start of sample_one_dimensional: the invalid-input and median cases, else the
// Newton start (the Gaussian quantile, corrected by the tabulated guess).
inline constexpr int kStarted = -1;  // Newton iteration to follow
template <bool Tab, int Cap>
C8MI_HD int newton_start(c8mi_moliere_snapshot const& m, Table const& t, State<Cap> const& st,
                         double u, double& current) {
  current = 0.;
  if (!isfinite(u) || !(u > 0.) || !(u < 1.)) return kInvalidInput;
  if (u == 0.5) return kSuccess;
  double const gaussian = normal_ppf(u) / kSqrt2;
  if constexpr (Tab) {
    double const reference_b = 1. / st.inverse_b[m.maximum_weight_index];
    current = st.prefactor * improve_guess(t.guess, reference_b, st.beta_squared, gaussian);
  } else {
    current = st.prefactor * gaussian;
  }
  return kStarted;
}

// One Newton update of sample_one_dimensional: `current` becomes the next
// iterate; kStarted while not converged.
C8MI_HD int newton_update(double density, double cdf, double target, double& current) {
  double const previous = current;
  if (!isfinite(density) || density == 0.) return kNonFinite;  // current stays the previous iterate
  current = previous - (cdf - target) / density;
  if (!isfinite(current)) return kNonFinite;
  if (fabs((previous - current) / current) <= 5.e-5) return kSuccess;
  return kStarted;
}

inline constexpr int kMaximumIterations = 100;
inline constexpr int kPending = -2;  // not converged within the iterations allowed here

// Newton iterations `first` .. `last` (1-based, at most 100) from `current`:
// kSuccess, kNonFinite, kNonConvergent after iteration 100, or kPending after
// `last` < 100.
template <bool Tab, int Cap>
C8MI_HD int newton_iterate(c8mi_moliere_snapshot const& m, Quotients const& q, Table const& t,
                           State<Cap> const& st, double u, int first, int last,
                           double& current) {
  double const target = u - 0.5;
  for (int it = first; it <= last; ++it) {
    double density, cdf;
    if constexpr (Tab) evaluate_table(m, q, t, st, current, density, cdf);
    else evaluate(m, q, st, current, density, cdf);
    int const status = newton_update(density, cdf, target, current);
    if (status != kStarted) return status;
  }
  return last >= kMaximumIterations ? kNonConvergent : kPending;
}

template <bool Tab, int Cap>
C8MI_HD int one_dimensional(c8mi_moliere_snapshot const& m, Quotients const& q, Table const& t,
                            State<Cap> const& st, double u, double& angle) {
  int const status = newton_start<Tab>(m, t, st, u, angle);
  if (status != kStarted) return status;
  return newton_iterate<Tab>(m, q, t, st, u, 1, kMaximumIterations, angle);
}

// ---- the slow tail: one wavefront per sample that did not converge early ----
//
// About one e+- step in 1e5 draws a far-tail quantile whose Newton iteration
// runs to its limit of 100 through the large-x expansions (two dozen pow per
// component and iteration). Serially that is ~1 ms of one lane, and the whole
// launch waits for it. The step kernels therefore stop after a few iterations
// (EmContext::moliere_fast_iterations); a sample still open continues here
// with its iterate: each evaluation spreads the transcendental parts of every
// component over a workgroup of kTailThreads (LDS `jobs`), then every thread
// combines them in the serial order, so density, CDF and iterates are
// bit-identical to the serial evaluation.
//
// Jobs per component: 13 pow calls (x^1.5 and (1 - 4.5/x)^(2/3) of f1's
// large-x form, the 11 terms of f2's expansion, each with its log), 11 pown
// calls (the terms of the expansion of f2's integral, each with its log), then
// exp(-x), erf(sqrt x) and f1's large-x integral. Each kind runs on a wave of
// its own (pow jobs on the first, pown jobs on the second, the other three on
// the third), so every wave runs one code path: a wave that mixed the kinds
// ran every kind's code one after another (divergence), ~2x the time.
inline constexpr int kPowJobs = 13;
inline constexpr int kPownJobs = 11;
inline constexpr int kJobSlots = kPowJobs + kPownJobs + 3;
inline constexpr int kTailThreads = 3 * 64;

// Horner of a 70-coefficient power series (the order of each chain of evaluate).
C8MI_HD double series70(double const* c, double x) {
  double s = c[69];
  for (int i = 68; i >= 0; --i) s = s * x + c[i];
  return s;
}

template <bool Tab, int N>
__device__ __forceinline__ void evaluate_wave(c8mi_moliere_snapshot const& m, Quotients const& q, Table const& t,
                                              State<N> const& st, double theta, double* jobs, double& density,
                                              double& cumulative) {
  double const t2 = theta * theta;
  int const nc = static_cast<int>(m.component_count);
  double const f1_from = Tab ? kTableUpper : 12.;
  auto scale_of = [&](int k) {  // st.inverse_scale[k] without a dynamic index into registers
    double scale = 0.;
#pragma unroll
    for (int c = 0; c < N; ++c)
      if (c == k) scale = st.inverse_scale[c];
    return scale;
  };
  __syncthreads();  // the previous evaluation has read `jobs`
  int const wave = static_cast<int>(threadIdx.x) / 64, lane = static_cast<int>(threadIdx.x) % 64;
  if (wave == 0) {
    // slot s: 0 x^1.5, 1 (1 - 4.5/x)^(2/3), 2..12 f2_term(i = s)
    if (lane < nc * kPowJobs) {
      int const k = lane / kPowJobs, s = lane % kPowJobs;
      double const x = t2 * scale_of(k);
      double const base = s == 1 ? 1. - 4.5 / x : x;
      double const p = pow(base, s == 0 ? 1.5 : s == 1 ? 2. / 3. : -(s + 0.5));
      double const term = m.c2_large[s] * (0.5 * log(x) + m.s2_large[s]) * p;
      jobs[k * kJobSlots + s] = s < 2 ? (x > f1_from ? p : 0.) : (x > 6.5 * 6.5 ? term : 0.);
    }
  } else if (wave == 1) {
    // slot kPowJobs + s: integral_f2_term(i = s + 2)
    if (lane < nc * kPownJobs) {
      int const k = lane / kPownJobs, i = lane % kPownJobs + 2;
      double const x = t2 * scale_of(k);
      jobs[k * kJobSlots + kPowJobs + i - 2] = x > 6.5 * 6.5 ? integral_f2_term(m, q, x, i) : 0.;
    }
  } else if (lane < nc * 3) {
    int const k = lane / 3, r = lane % 3;
    double const x = t2 * scale_of(k);
    double v;
    if (r == 0) v = exp(-x);
    else if (r == 1) v = erf(sqrt(x > 0. ? x : 0.));
    else v = x > f1_from ? integral_f1_large(m, x) : 0.;
    jobs[k * kJobSlots + kPowJobs + kPownJobs + r] = v;
  }
  __syncthreads();
  double dens = 0., cdf = 0.;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    if (k >= nc) break;
    double const* J = jobs + k * kJobSlots;
    double const x = t2 * st.inverse_scale[k], ib = st.inverse_b[k], ib2 = ib * ib;
    double const root = sqrt(x > 0. ? x : 0.);
    double f1v, f2v, i1v, i2v;
    // f1_large's expression over its two pow jobs
    auto f1_above = [&] { return 0.5 * sqrt(kPi) / (J[0] * J[1]); };
    auto f2_above = [&] {
      if (x <= 6.5 * 6.5) return f2_fit(x);
      double sum = 0.;
      for (int i = 2; i < 13; ++i) sum += J[i];
      return sum;
    };
    auto i2_above = [&] {
      if (x <= 6.5 * 6.5) return integral_f2_fit(x);
      double sum = 0.;
      for (int i = 2; i < 13; ++i) sum += J[kPowJobs + i - 2];
      return sum;
    };
    if (Tab && x <= kTableUpper) {
      double coordinate = (x - 0.) * (static_cast<double>(kIntervals) / (kTableUpper - 0.));
      if (coordinate < 0.) coordinate = 0.;
      if (coordinate > static_cast<double>(kIntervals)) coordinate = static_cast<double>(kIntervals);
      auto interval = static_cast<std::uint32_t>(coordinate);
      if (interval >= static_cast<std::uint32_t>(kIntervals)) interval = kIntervals - 1;
      double const u = coordinate - static_cast<double>(interval);
      auto cubic = [&](int f) {
        double const* c = t.polynomials[f * kIntervals + interval].c;
        return c[0] + u * (c[1] + u * (c[2] + u * c[3]));
      };
      f1v = cubic(0);
      f2v = cubic(1);
      i1v = cubic(2);
      i2v = cubic(3);
    } else if (Tab) {
      f1v = f1_above();
      i1v = J[kPowJobs + kPownJobs + 2];
      f2v = f2_above();
      i2v = i2_above();
    } else {
      f1v = x > 12. ? f1_above() : series70(m.c1, x);
      i1v = x > 12. ? J[kPowJobs + kPownJobs + 2] : series70(q.i1, x) * root;
      f2v = x > 4.25 * 4.25 ? f2_above() : series70(m.c2, x);
      i2v = x > 4.25 * 4.25 ? i2_above() : series70(q.i2, x) * root;
    }
    dens += st.density_weight[k] * (J[kPowJobs + kPownJobs] + f1v * ib + f2v * ib2);
    cdf += m.components[k].weight_zz * (0.5 * J[kPowJobs + kPownJobs + 1] + sqrt(1. / kPi) * (i1v * ib + i2v * ib2));
  }
  density = dens * m.inverse_weight_zz_sum;
  cdf *= m.inverse_weight_zz_sum;
  cumulative = theta < 0. ? -cdf : cdf;
}

// newton_iterate with the workgroup evaluation (every thread the same iterates).
template <bool Tab, int Cap>
__device__ __forceinline__ int newton_iterate_wave(c8mi_moliere_snapshot const& m, Quotients const& q,
                                                   Table const& t, State<Cap> const& st, double u, int first,
                                                   double* jobs, double& current) {
  double const target = u - 0.5;
  for (int it = first; it <= kMaximumIterations; ++it) {
    double density, cdf;
    evaluate_wave<Tab>(m, q, t, st, current, jobs, density, cdf);
    int const status = newton_update(density, cdf, target, current);
    if (status != kStarted) return status;
  }
  return kNonConvergent;
}

// Space angle of two projected draws (sample_moliere_scattering_angle_2d).
template <bool Tab, int Cap>
C8MI_HD int angle_2d_for(c8mi_moliere_snapshot const& m, Quotients const& q, Table const& t,
                         double grammage, double e_mev, double u1, double u2, double& angle) {
  State<Cap> st;
  angle = 0.;
  int status = build(m, grammage, e_mev, st);
  if (status != kSuccess) return status;
  double a1, a2;
  status = one_dimensional<Tab>(m, q, t, st, u1, a1);
  if (status != kSuccess) return status;
  status = one_dimensional<Tab>(m, q, t, st, u2, a2);
  if (status != kSuccess) return status;
  angle = sqrt(a1 * a1 + a2 * a2);
  return isfinite(angle) ? kSuccess : kNonFinite;
}

// Media of up to kDeviceComponents components (air has three; the backend
// rejects larger Molière snapshots at create).
inline constexpr int kDeviceComponents = 4;
static_assert(kDeviceComponents * kPowJobs <= 64 && kDeviceComponents * kPownJobs <= 64 && kDeviceComponents * 3 <= 64);

// apply_moliere_direction_cpu_step_compatible: normalize(tracked + rotate(start) - start).
C8MI_HD int deflect_end(double const start[3], double const tracked[3], double angle, double au,
                        double out[3]) {
  if (!isfinite(angle) || angle < 0. || !isfinite(au) || !(au > 0.) || !(au < 1.)) return kInvalidInput;
  constexpr double lo = (1. - 1.e-12) * (1. - 1.e-12), hi = (1. + 1.e-12) * (1. + 1.e-12);
  double const n2 = start[0] * start[0] + start[1] * start[1] + start[2] * start[2];
  if (!isfinite(n2) || n2 < lo || n2 > hi) return kInvalidInput;
  double a1[3] = {0., 0., 0.};
  if (start[0] > 0.1) {
    a1[0] = -start[1];
    a1[1] = start[0];
  } else {
    a1[1] = -start[2];
    a1[2] = start[1];
  }
  double const an = sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
  if (!isfinite(an) || !(an > 0.)) return kInvalidInput;
  for (int k = 0; k < 3; ++k) a1[k] /= an;
  double const ax[3] = {a1[1] * start[2] - a1[2] * start[1], a1[2] * start[0] - a1[0] * start[2],
                        a1[0] * start[1] - a1[1] * start[0]};
  double const ct = cos(angle), mst = -sin(angle);
  double z[3];
  for (int k = 0; k < 3; ++k) z[k] = start[k] * ct + ax[k] * mst;
  double const az = au * 2. * kPi;
  double const ca = cos(az), msa = -sin(az);
  double const dz[3] = {start[1] * z[2] - start[2] * z[1], start[2] * z[0] - start[0] * z[2],
                        start[0] * z[1] - start[1] * z[0]};
  double const par = start[0] * z[0] + start[1] * z[1] + start[2] * z[2];
  double sc[3];
  for (int k = 0; k < 3; ++k) sc[k] = z[k] * ca + dz[k] * msa + start[k] * par * (1. - ca);
  double const sn = sqrt(sc[0] * sc[0] + sc[1] * sc[1] + sc[2] * sc[2]);
  if (!isfinite(sn) || !(sn > 0.)) return kNonFinite;
  for (int k = 0; k < 3; ++k) sc[k] /= sn;
  double const t2 = tracked[0] * tracked[0] + tracked[1] * tracked[1] + tracked[2] * tracked[2];
  if (!isfinite(t2) || t2 < lo || t2 > hi) return kInvalidInput;
  double norm2 = 0.;
  for (int k = 0; k < 3; ++k) {
    out[k] = tracked[k] + sc[k] - start[k];
    norm2 += out[k] * out[k];
  }
  double const norm = sqrt(norm2);
  if (!isfinite(norm) || !(norm > 0.)) return kNonFinite;
  for (int k = 0; k < 3; ++k) out[k] /= norm;
  return kSuccess;
}

}  // namespace c8mi::em::moliere
