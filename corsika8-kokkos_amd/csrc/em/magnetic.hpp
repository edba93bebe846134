// e+- transport in a uniform magnetic field on the device: the leapfrog step
// limit and advance, and the boundary and observation-plane intersections of
// the curved (leapfrog) trajectory.
//
// Reference: corsika/detail/kokkos/environment/UniformMagneticField.inl
// (maximum_uniform_magnetic_step, advance_uniform_magnetic_field),
// EnvironmentLeptonBoundary.inl (lepton_boundary, lepton_observation_distance),
// CurvedBoundaryAccess.inl:38-126 (intersectShape), and the root finders they
// use: corsika/detail/kokkos/utility/{PolynomialRoots,StableQuartic,
// ExtendedPrecision}.inl. The reference's device build solves the quadratic
// and quartic in double-double arithmetic (its host build in long double);
// this file follows the device build's operation sequences so device and
// oracle give identical roots.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "em/environment.hpp"

namespace c8mi::em::mag {

enum Status : int { kSuccess = 0, kLinear = 1, kInvalidInput = 2, kNonFinite = 3 };

inline constexpr double kSpeedOfLight = 299792458.;
inline constexpr double kChargeC = 1.6021766208e-19;
inline constexpr double kRigidityGeVPerTm = 0.2997925146834389;

// ---- double-double numbers (value = hi + lo, |lo| <= ulp(hi) / 2) ----

struct DD {
  double hi, lo;
};

C8MI_HD DD dd(double v) { return {v, 0.}; }
C8MI_HD double to_double(DD a) { return a.hi + a.lo; }
// Fast two-sum of a and b, |a| >= |b| or a = 0.
C8MI_HD DD renormalize(double a, double b) {
  double const s = a + b;
  return {s, b - (s - a)};
}
C8MI_HD DD operator+(DD a, DD b) {
  double const s = a.hi + b.hi;
  double const bv = s - a.hi;
  double const e = (a.hi - (s - bv)) + (b.hi - bv);
  return renormalize(s, e + (a.lo + b.lo));
}
C8MI_HD DD operator-(DD a) { return {-a.hi, -a.lo}; }
C8MI_HD DD operator-(DD a, DD b) { return a + (-b); }
C8MI_HD DD operator*(DD a, DD b) {
  double const p = a.hi * b.hi;
  double e = fma(a.hi, b.hi, -p);  // the exact rounding error of p
  e = e + (a.hi * b.lo + a.lo * b.hi);
  e = e + a.lo * b.lo;
  return renormalize(p, e);
}
// Long division with two correction terms.
C8MI_HD DD operator/(DD a, DD b) {
  double const q = a.hi / b.hi;
  if (!isfinite(q)) return dd(q);
  DD r = a - b * dd(q);
  double const c = r.hi / b.hi;
  r = r - b * dd(c);
  return (dd(q) + dd(c)) + dd(r.hi / b.hi);
}
C8MI_HD bool operator<(DD a, DD b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }
C8MI_HD bool operator>(DD a, DD b) { return b < a; }
C8MI_HD bool operator==(DD a, DD b) { return a.hi == b.hi && a.lo == b.lo; }
C8MI_HD DD dd_abs(DD a) { return a < dd(0.) ? -a : a; }
// One Newton correction of the double square root.
C8MI_HD DD dd_sqrt(DD v) {
  double const a = sqrt(v.hi);
  if (a == 0. || !isfinite(a)) return dd(a);
  return dd(a) + (v - dd(a) * dd(a)) / dd(2. * a);
}
// Two Newton corrections of the double cube root.
C8MI_HD DD dd_cbrt(DD v) {
  DD a = dd(cbrt(v.hi));
  if (a.hi == 0. || !isfinite(a.hi)) return a;
  for (int it = 0; it < 2; ++it) a = a + (v - a * a * a) / (dd(3.) * a * a);
  return a;
}
C8MI_HD DD dd_acos(DD v) {
  double const angle = acos(v.hi);
  if (v.lo == 0.) return dd(angle);
  if (v.hi == 1.) return dd_sqrt(dd(2.) * (dd(1.) - v));
  if (v.hi == -1.) return dd(3.14159265358979323846) - dd_sqrt(dd(2.) * (dd(1.) + v));
  return dd(angle) - dd(v.lo / sqrt((1. - v.hi) * (1. + v.hi)));
}
C8MI_HD DD dd_cos(DD v) { return dd(cos(v.hi)) - dd(sin(v.hi) * v.lo); }
// fma of the double-double type: a product and a sum, each double-double.
C8MI_HD DD dd_fma(DD a, DD b, DD c) { return a * b + c; }

// ---- real roots (at most four, as doubles) ----
// The general solvers stay out of line: they run only near a boundary in a
// field, and inlined they enlarged the e+- step kernels' frames.

struct Roots {
  double v[4];
  unsigned n;
};

C8MI_HD void append(Roots& r, double x) { r.v[r.n++] = x; }
// Skips a value within 8 ulp-scale of one already present; keeps at most four.
C8MI_HD void append_distinct(Roots& r, double x) {
  for (unsigned i = 0; i < r.n; ++i)
    if (fabs(x - r.v[i]) <= 8 * 2.220446049250313e-16 * fmax(fabs(x), fabs(r.v[i]))) return;
  if (r.n < 4) append(r, x);
}

// a x^2 + b x + c = 0 (PolynomialRoots.inl:104-134): the linear cases when a or
// c is below epsilon (inline: the observation plane's SI coefficients always
// take the first), else the cancellation-free pair (out of line).
C8MI_HD_NOINLINE Roots quadratic_pair(DD a, DD b, DD c, double epsilon) {
  Roots r{{0., 0., 0., 0.}, 0u};
  DD const multiplied = a * dd(4.) * c;
  DD const error = dd_fma(a * dd(4.), c, -multiplied);
  DD const corrected = dd_fma(b, b, -multiplied);
  DD const radicand = corrected + error;
  if (dd_abs(radicand) < dd(epsilon)) {
    append(r, to_double(-b / (dd(2.) * a)));
    return r;
  }
  if (radicand < dd(0.)) return r;
  DD const root = c * dd(2.) / (b > dd(0.) ? -b - dd_sqrt(radicand) : -b + dd_sqrt(radicand));
  append(r, to_double(root));
  append(r, to_double(c / (a * root)));
  return r;
}

C8MI_HD Roots quadratic_roots(DD a, DD b, DD c, double epsilon) {
  Roots r{{0., 0., 0., 0.}, 0u};
  if (dd_abs(a) < dd(epsilon)) {
    double const bb = to_double(b), cc = to_double(c);
    if (bb != 0.) append(r, -cc / bb);
    return r;
  }
  if (dd_abs(c) < dd(epsilon)) {
    double const aa = to_double(a), bb = to_double(b);
    if (aa != 0.) append(r, -bb / aa);
    append(r, 0.);
    return r;
  }
  return quadratic_pair(a, b, c, epsilon);
}

// x^2 + b x + c = 0 without cancellation (StableQuartic.inl stable_monic_quadratic).
C8MI_HD Roots monic_quadratic(DD b, DD c) {
  Roots r{{0., 0., 0., 0.}, 0u};
  if (c == dd(0.)) {
    append(r, 0.);
    if (!(b == dd(0.))) append(r, to_double(-b));
    return r;
  }
  DD const multiplied = dd(4.) * c;
  DD const disc = dd_fma(b, b, -multiplied);
  if (disc < dd(0.)) return r;
  if (disc == dd(0.)) {
    append(r, to_double(-b / dd(2.)));
    return r;
  }
  DD const large = -(b + (!(b < dd(0.)) ? dd_sqrt(disc) : -dd_sqrt(disc))) / dd(2.);
  append(r, to_double(large));
  append(r, to_double(c / large));
  return r;
}

C8MI_HD DD halley(DD root, DD p, DD quadratic, DD negative_constant) {
  for (int it = 0; it < 4; ++it) {
    DD const value = ((root + p) * root + quadratic) * root - negative_constant;
    DD const derivative = (dd(3.) * root + dd(2.) * p) * root + quadratic;
    DD const second = dd(6.) * root + dd(2.) * p;
    root = root - dd(2.) * value * derivative / (dd(2.) * derivative * derivative - value * second);
  }
  return root;
}

// The positive root of Ferrari's resolvent cubic of x^4 + p x^2 + q x + r
// (stable_resolvent): a Halley-refined small root where it is the only positive
// one, else Cardano or the trigonometric form.
C8MI_HD DD resolvent(DD p, DD q, DD r) {
  DD const quadratic = p * p / dd(4.) - r;
  DD const negative_constant = q * q / dd(8.);
  if (p < dd(0.) && r < dd(0.) && negative_constant < (-p) * quadratic / dd(64.) &&
      negative_constant < p * r / dd(8.))
    return halley(negative_constant / quadratic, p, quadratic, negative_constant);
  if (!(p < dd(0.)) && !(quadratic < dd(0.))) {
    DD root = dd_cbrt(negative_constant);
    if (p > dd(0.)) {
      DD const candidate = dd_sqrt(negative_constant / p);
      if (candidate < root) root = candidate;
    }
    if (quadratic > dd(0.)) {
      DD const candidate = negative_constant / quadratic;
      if (candidate < root) root = candidate;
    }
    return halley(root, p, quadratic, negative_constant);
  }
  DD const dq = (dd(3.) * quadratic - p * p) / dd(9.);
  DD const dc = (p * (dd(9.) * quadratic - dd(2.) * p * p) + dd(27.) * negative_constant) / dd(54.);
  DD const disc = dd_fma(dc, dc, dq * dq * dq);
  if (!(disc < dd(0.))) {
    DD const cube = dd_cbrt(dc + (!(dc < dd(0.)) ? dd_sqrt(disc) : -dd_sqrt(disc)));
    return (cube == dd(0.) ? dd(0.) : cube - dq / cube) - p / dd(3.);
  }
  DD argument = dc / dd_sqrt(-dq * dq * dq);
  if (argument > dd(1.)) argument = dd(1.);
  if (argument < dd(-1.)) argument = dd(-1.);
  return dd(2.) * dd_sqrt(-dq) * dd_cos(dd_acos(argument) / dd(3.)) - p / dd(3.);
}

// Real roots of x^4 + p x^2 + q x + r, ascending (stable_quartic_roots): scaled
// by a power of two, biquadratic when q = 0, else Ferrari's factorisation into
// two monic quadratics; a non-positive resolvent yields the single root +inf
// (the caller's incomplete-roots signal).
C8MI_HD_NOINLINE Roots quartic_roots(double p, double q, double r) {
  Roots out{{0., 0., 0., 0.}, 0u};
  int pe = -4096, qe = -4096, re = -4096;
  if (p != 0.) {
    int e;
    frexp(p, &e);
    pe = e >= 0 ? (e + 1) / 2 : e / 2;
  }
  if (q != 0.) {
    int e;
    frexp(q, &e);
    qe = e >= 0 ? (e + 2) / 3 : e / 3;
  }
  if (r != 0.) {
    int e;
    frexp(r, &e);
    re = e >= 0 ? (e + 3) / 4 : e / 4;
  }
  int exponent = pe > qe ? pe : qe;
  exponent = exponent > re ? exponent : re;
  if (exponent == -4096) {
    append(out, 0.);
    return out;
  }
  DD const sp = dd(ldexp(p, -2 * exponent)), sq = dd(ldexp(q, -3 * exponent)), sr = dd(ldexp(r, -4 * exponent));
  double const scale = ldexp(1., exponent);
  if (q == 0.) {
    Roots const squares = monic_quadratic(sp, sr);
    for (unsigned i = 0; i < squares.n; ++i) {
      if (squares.v[i] < 0.) continue;
      if (squares.v[i] == 0.) {
        append_distinct(out, 0.);
      } else {
        double const root = to_double(dd_sqrt(dd(squares.v[i]))) * scale;
        append_distinct(out, -root);
        append_distinct(out, root);
      }
    }
  } else {
    DD const y = resolvent(sp, sq, sr);
    if (!(y > dd(0.))) {
      append(out, __builtin_inf());
      return out;
    }
    DD const linear = dd_sqrt(dd(2.) * y);
    DD const common = sp / dd(2.) + y;
    DD const difference = sq / (dd(2.) * linear);
    DD first, second;
    // the two constants multiply to r: the larger without cancellation, the other by division
    if (!(dd_abs(common + difference) < dd_abs(common - difference))) {
      second = common + difference;
      first = sr / second;
    } else {
      first = common - difference;
      second = sr / first;
    }
    Roots const a = monic_quadratic(linear, first);
    Roots const b = monic_quadratic(-linear, second);
    for (unsigned i = 0; i < a.n; ++i) append_distinct(out, a.v[i] * scale);
    for (unsigned i = 0; i < b.n; ++i) append_distinct(out, b.v[i] * scale);
  }
  for (unsigned i = 1; i < out.n; ++i) {  // ascending
    double const v = out.v[i];
    unsigned j = i;
    while (j && v < out.v[j - 1]) {
      out.v[j] = out.v[j - 1];
      --j;
    }
    out.v[j] = v;
  }
  return out;
}

// ---- step limit and advance ----

struct Limit {
  int status;
  double distance_m;
  double gyroradius_m;
};

C8MI_HD bool valid_particle(c8mi_particle_state const& p, double mass_gev, double charge) {
  if (!isfinite(p.energy_gev) || !(p.energy_gev > mass_gev) || !isfinite(mass_gev) || !(mass_gev > 0.) ||
      !isfinite(charge))
    return false;
  for (int a = 0; a < 3; ++a)
    if (!isfinite(p.position_m[a]) || !isfinite(p.direction[a])) return false;
  double const n2 = env::dot3(p.direction, p.direction);
  return isfinite(n2) && fabs(n2 - 1.) <= 1.e-12;
}

C8MI_HD bool valid_field(double const b[3]) {
  return isfinite(b[0]) && isfinite(b[1]) && isfinite(b[2]);
}

// maximum_uniform_magnetic_step: the leapfrog length that turns the direction
// by at most `max_deflection` on the gyro circle, 2 cos(d) sin(d) R; Linear
// (infinite) for neutral particles, no field, p_perp < 1 eV or R > 1e9 m.
C8MI_HD Limit step_limit(c8mi_particle_state const& p, double mass_gev, double charge,
                         double const b[3], double max_deflection) {
  if (!valid_particle(p, mass_gev, charge) || !valid_field(b) || !isfinite(max_deflection) ||
      !(max_deflection > 0.) || !(max_deflection < 0.5 * 3.141592653589793))
    return {kInvalidInput, 0., 0.};
  double const b2 = env::dot3(b, b);
  double const inf = __builtin_inf();
  if (charge == 0. || b2 == 0.) return {kLinear, inf, inf};
  double const momentum = sqrt((p.energy_gev - mass_gev) * (p.energy_gev + mass_gev));
  double const pv[3] = {momentum * p.direction[0], momentum * p.direction[1], momentum * p.direction[2]};
  double const along = env::dot3(pv, b) / b2;  // the vector projection, not a difference of squares
  double const perp[3] = {pv[0] - along * b[0], pv[1] - along * b[1], pv[2] - along * b[2]};
  double const p_perp = sqrt(env::dot3(perp, perp));
  if (p_perp < 1.e-9) return {kLinear, inf, inf};
  double const radius = p_perp / (kRigidityGeVPerTm * fabs(charge) * sqrt(b2));
  if (!isfinite(radius) || radius > 1.e9) return {kLinear, inf, radius};
  double const distance = 2. * cos(max_deflection) * sin(max_deflection) * radius;
  if (!isfinite(distance) || !(distance > 0.)) return {kNonFinite, 0., radius};
  return {kSuccess, distance, radius};
}

// advance_uniform_magnetic_field: one leapfrog step of length `distance`: the
// direction gets the full kick (q/p) L (u x B), the position moves along the
// half-kicked direction; the time advances by L / (beta c).
C8MI_HD int advance(c8mi_particle_state const& start, double mass_gev, double charge,
                    double const b[3], double distance, c8mi_particle_state& end) {
  if (!valid_particle(start, mass_gev, charge) || !valid_field(b) || !isfinite(distance) || distance < 0.)
    return kInvalidInput;
  end = start;
  if (distance == 0.) return kSuccess;
  double const momentum = sqrt((start.energy_gev - mass_gev) * (start.energy_gev + mass_gev));
  double const coefficient = charge * kRigidityGeVPerTm * distance / momentum;
  double const uxb[3] = {start.direction[1] * b[2] - start.direction[2] * b[1],
                         start.direction[2] * b[0] - start.direction[0] * b[2],
                         start.direction[0] * b[1] - start.direction[1] * b[0]};
  double kicked[3], disp[3];
  for (int a = 0; a < 3; ++a) {
    double const kick = coefficient * uxb[a];
    kicked[a] = start.direction[a] + kick;
    disp[a] = distance * (start.direction[a] + 0.5 * kick);
    end.position_m[a] = start.position_m[a] + disp[a];
  }
  double const kn = sqrt(env::dot3(kicked, kicked));
  double const chord = sqrt(env::dot3(disp, disp));
  double const beta = momentum / start.energy_gev;
  if (!isfinite(kn) || !(kn > 0.) || !isfinite(chord) || !isfinite(beta) || !(beta > 0.)) return kNonFinite;
  for (int a = 0; a < 3; ++a) end.direction[a] = kicked[a] / kn;
  end.time_s = start.time_s + distance / (beta * kSpeedOfLight);
  return env::dot3(uxb, uxb) == 0. || charge == 0. ? kLinear : kSuccess;
}

// ---- curved-trajectory intersections (CurvedBoundaryAccess) ----

struct Curved {
  double position[3], velocity[3], direction[3], field[3];
  double speed, momentum_si, charge_si, gyroradius;
  bool linear;
  mutable bool unsupported, incomplete;
};

C8MI_HD Curved make_curved(c8mi_particle_state const& p, double mass_gev, double charge,
                           double const b[3], Limit const& limit, bool linear) {
  Curved a;
  double const momentum = sqrt((p.energy_gev - mass_gev) * (p.energy_gev + mass_gev));
  a.speed = kSpeedOfLight * momentum / p.energy_gev;
  a.charge_si = charge * kChargeC;
  a.momentum_si = momentum * kChargeC / kRigidityGeVPerTm;
  for (int k = 0; k < 3; ++k) {
    a.position[k] = p.position_m[k];
    a.velocity[k] = a.speed * p.direction[k];
    a.direction[k] = p.direction[k];
    a.field[k] = b[k];
  }
  a.gyroradius = limit.gyroradius_m;
  a.linear = linear;
  a.unsupported = a.incomplete = false;
  return a;
}

// Time parameters of the entry and exit of a sphere (or the universe), or of a
// plane (center/normal) when `plane` is set.
C8MI_HD env::Intersections curved_intersect(Curved const& a, std::uint32_t shape_kind,
                                            double const center[3], double radius,
                                            double const normal[3], bool plane) {
  env::Intersections const none{0., 0., false};
  double const inf = __builtin_inf();
  auto straight = [&] {
    c8mi_env_node s{};
    s.shape_kind = shape_kind;
    for (int k = 0; k < 3; ++k) s.center_m[k] = center[k];
    s.radius_m = radius;
    return env::intersect(s, a.position, a.velocity);
  };
  if (!plane && a.linear) return straight();
  if (!plane && shape_kind == C8MI_SHAPE_UNIVERSE) return none;
  double const projected[3] = {a.direction[1] * a.field[2] - a.direction[2] * a.field[1],
                               a.direction[2] * a.field[0] - a.direction[0] * a.field[2],
                               a.direction[0] * a.field[1] - a.direction[1] * a.field[0]};
  double const delta[3] = {a.position[0] - center[0], a.position[1] - center[1], a.position[2] - center[2]};
  if (plane) {
    // e+- carry charge, so the straight plane branch of a neutral particle is not needed
    double const denominator = a.charge_si * env::dot3(normal, projected);
    double const linear = 2 * a.momentum_si * env::dot3(a.direction, normal);
    double const constant = 2 * a.momentum_si * env::dot3(normal, delta);
    Roots const r = quadratic_roots(dd(denominator), dd(linear), dd(constant), 1.e-12);
    double first = inf;
    for (unsigned i = 0; i < r.n; ++i)
      if (r.v[i] >= 0. && r.v[i] < first) first = r.v[i];
    return {first / a.speed, inf, true};
  }
  double const projected2 = env::dot3(projected, projected);
  if (projected2 == 0.) return straight();
  double const delta_norm = sqrt(env::dot3(delta, delta));
  double const distance = fabs(delta_norm - radius);
  if (distance > a.gyroradius) return none;
  if (a.gyroradius > 1000 * distance) return straight();
  double const k = a.charge_si / a.momentum_si;
  double const denominator = 4. / (projected2 * k * k);
  double const quadratic = (env::dot3(projected, delta) * k + 1) * denominator;
  double const linear = env::dot3(a.direction, delta) * 2 * denominator;
  double const constant = (delta_norm + radius) * (delta_norm - radius) * denominator;
  Roots const roots = quartic_roots(quadratic, linear, constant);
  bool const inside = env::dot3(delta, delta) < radius * radius;
  double entry = 0., exit = 0.;
  unsigned count = 0, entries = 0, exits = 0;
  for (unsigned i = 0; i < roots.n; ++i) {
    double const d = roots.v[i];
    if (!isfinite(d)) {
      a.incomplete = true;
      return none;
    }
    if (inside) {
      if (d < 1.e-4) {
        if (!entries || entry < d) entry = d;
        ++entries;
      } else {
        if (!exits || d < exit) exit = d;
        ++exits;
      }
      count = (entries > 0) + (exits > 0);
    } else {
      if (d < -1.e-4) continue;
      if (!count) {
        entry = d;
      } else if (d < entry) {
        exit = entry;
        entry = d;
      } else {
        exit = d;
      }
      ++count;
    }
  }
  if (!count) return none;
  if ((inside && (!entries || !exits)) || (!inside && count < 2)) {  // an unresolved boundary
    a.incomplete = true;
    return none;
  }
  return {entry / a.speed, exit / a.speed, true};
}

C8MI_HD env::Intersections curved_node(EmContext const& c, Curved const& a, std::int32_t node) {
  c8mi_env_node const& n = c.nodes[node];
  double const zero[3] = {0., 0., 0.};
  return curved_intersect(a, n.shape_kind, n.center_m, n.radius_m, zero, false);
}

// lepton_boundary: next_environment_boundary over the curved access with the
// magnetic limit as the parameter bound; distance, next node and failure.
C8MI_HD bool lepton_boundary(EmContext const& c, c8mi_particle_state const& p, double mass_gev,
                             double charge, double const b[3], Limit const& limit,
                             double& distance, std::int32_t& next) {
  Curved const a = make_curved(p, mass_gev, charge, b, limit, limit.status == kLinear);
  std::int32_t const current = p.environment_node;
  double selected = limit.distance_m / a.speed;
  std::int32_t selected_node = current;
  c8mi_env_node const& n = c.nodes[current];
  env::Intersections const own = curved_node(c, a, current);
  if (own.has && own.exit <= selected) {
    selected = own.exit;
    selected_node = n.parent;
  }
  for (std::uint32_t k = 0; k < n.child_count; ++k) {
    std::int32_t const child = c.children[n.child_begin + k];
    env::Intersections const x = curved_node(c, a, child);
    if (!x.has) continue;
    if (x.exit > 0. && x.entry <= selected && -x.entry < x.exit) {
      selected = x.entry;
      selected_node = child;
    }
  }
  if (selected < 0.) selected = 0.;
  distance = selected * a.speed;
  next = selected_node;
  return !(a.unsupported || a.incomplete);
}

// lepton_observation_distance: the curved trajectory always (also where the
// boundary search ran straight); +inf when the plane is behind or missed.
C8MI_HD double lepton_observation_distance(EmContext const& c, c8mi_particle_state const& p,
                                           double mass_gev, double charge, double const b[3],
                                           Limit const& limit) {
  Curved const a = make_curved(p, mass_gev, charge, b, limit, false);
  env::Intersections const x =
      curved_intersect(a, 0u, c.observation_point, 0., c.observation_normal, true);
  return x.entry > 0. ? x.entry * a.speed : __builtin_inf();
}

}  // namespace c8mi::em::mag
