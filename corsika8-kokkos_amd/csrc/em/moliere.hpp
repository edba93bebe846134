// Molière multiple scattering of one e+- step on the device: the projected-angle
// distribution with the analytical correction functions, two Newton-solved
// projected angles, and the deflection of the step's end direction.
//
// Reference: corsika/detail/kokkos/em/models/MoliereScattering.inl:219-688
// (normal_ppf, solve_moliere_b, f1_m / f2_m and their integrals, build_distribution,
// evaluate_distribution, sample_one_dimensional, apply_moliere_direction and its
// step-compatible form) and corsika/detail/kokkos/em/algorithms/MoliereStep.inl:11-101.
// The snapshot (c8mi_moliere_snapshot) sits in the EmContext and is read through
// a uniform pointer, so its coefficients are scalar loads shared by the wavefront.
#pragma once

#include <hip/hip_runtime.h>

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
__device__ __forceinline__ double normal_ppf(double p) {
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
__device__ __forceinline__ double solve_b(double rhs) {
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

// The snapshot was validated when the backend was created.
template <int Cap>
__device__ __forceinline__ int build(c8mi_moliere_snapshot const& m, double grammage, double e_mev, State<Cap>& st) {
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
__device__ __noinline__ double f1_large(double x) {
  return 0.5 * sqrt(kPi) / (pow(x, 1.5) * pow(1. - 4.5 / x, 2. / 3.));
}
__device__ __noinline__ double integral_f1_large(c8mi_moliere_snapshot const& m, double x) {
  double sum = m.c1_integral_large[14];
  for (int i = 13; i >= 0; --i) sum = m.c1_integral_large[i] + sum / x;
  return sum;
}
__device__ __noinline__ double f2_large(c8mi_moliere_snapshot const& m, double x) {
  if (x <= 6.5 * 6.5)
    return 0.00013567765224589459194769192063035 * x - 0.0022635502525409950842771866774683 * sqrt(x) +
           0.0098037758070269476889935233998585;
  double sum = 0.;
  for (int i = 2; i < 13; ++i) sum += m.c2_large[i] * (0.5 * log(x) + m.s2_large[i]) * pow(x, -(i + 0.5));
  return sum;
}
__device__ __noinline__ double integral_f2_large(c8mi_moliere_snapshot const& m, Quotients const& q, double x) {
  if (x <= 6.5 * 6.5)
    return -0.00026360133958801203364619158975302 * x + 0.0039965027465457608410459577896745 * sqrt(x) -
           0.016305842044996649714549974419242;
  double sum = 0.;
  for (int i = 2; i < 13; ++i) sum += q.i2_large[i] * (q.half_over[i] + 0.5 * log(x) + m.s2_large[i]) * pow(x, -i);
  return sum;
}

// Density and CDF of one projected angle (evaluate_distribution). The four
// power series of all N components run in one loop over the coefficients:
// 4N independent Horner chains share each coefficient load and fill the FP64
// pipeline, each chain with the same operations in the same order as alone.
template <int N>
__device__ __forceinline__ void evaluate(c8mi_moliere_snapshot const& m, Quotients const& q, State<N> const& st,
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

template <int Cap>
__device__ __forceinline__ int one_dimensional(c8mi_moliere_snapshot const& m, Quotients const& q,
                                               State<Cap> const& st, double u, double& angle) {
  angle = 0.;
  if (!isfinite(u) || !(u > 0.) || !(u < 1.)) return kInvalidInput;
  if (u == 0.5) return kSuccess;
  double current = st.prefactor * (normal_ppf(u) / kSqrt2);
  double const target = u - 0.5;
  for (int it = 1; it <= 100; ++it) {
    double const previous = current;
    double density, cdf;
    evaluate(m, q, st, previous, density, cdf);
    if (!isfinite(density) || density == 0.) {
      angle = previous;
      return kNonFinite;
    }
    current = previous - (cdf - target) / density;
    angle = current;
    if (!isfinite(current)) return kNonFinite;
    if (fabs((previous - current) / current) <= 5.e-5) return kSuccess;
  }
  return kNonConvergent;
}

// Space angle of two projected draws (sample_moliere_scattering_angle_2d).
template <int Cap>
__device__ __forceinline__ int angle_2d_for(c8mi_moliere_snapshot const& m, Quotients const& q, double grammage,
                                            double e_mev, double u1, double u2, double& angle) {
  State<Cap> st;
  angle = 0.;
  int status = build(m, grammage, e_mev, st);
  if (status != kSuccess) return status;
  double a1, a2;
  status = one_dimensional(m, q, st, u1, a1);
  if (status != kSuccess) return status;
  status = one_dimensional(m, q, st, u2, a2);
  if (status != kSuccess) return status;
  angle = sqrt(a1 * a1 + a2 * a2);
  return isfinite(angle) ? kSuccess : kNonFinite;
}

// Media of up to kDeviceComponents components (air has three; the backend
// rejects larger Molière snapshots at create).
inline constexpr int kDeviceComponents = 4;
__device__ __forceinline__ int angle_2d(c8mi_moliere_snapshot const& m, Quotients const& q, double grammage,
                                     double e_mev, double u1, double u2, double& angle) {
  return angle_2d_for<kDeviceComponents>(m, q, grammage, e_mev, u1, u2, angle);
}

// apply_moliere_direction_cpu_step_compatible: normalize(tracked + rotate(start) - start).
__device__ __forceinline__ int deflect_end(double const start[3], double const tracked[3], double angle, double au,
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
