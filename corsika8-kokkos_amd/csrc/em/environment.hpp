// Device environment queries of the EM step: a volume tree of spheres (and
// the universe) with exponential or homogeneous media, and the observation
// plane. Follows corsika/detail/kokkos/environment/{EnvironmentQuery,
// ModelEvaluation,ModelGrammage,StraightIntersections,EnvironmentSearch,
// ObservationPlane,EnvironmentLeptonBoundary}.inl and
// corsika/detail/framework/geometry/NextEnvironmentBoundaryBody.inl, with the
// separate roundings of their dot products (the library is built without
// floating-point contraction).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "em/em_device.hpp"

namespace c8mi::em::env {

enum Status : int { kSuccess = 0, kInvalid = 1, kNonFinite = 2, kOutside = 3, kNoForward = 5, kGrammageOutOfRange = 6 };

C8MI_HD double dot3(double const a[3], double const b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

C8MI_HD bool normalized(double const d[3]) {
  double const n = dot3(d, d);
  return isfinite(n) && fabs(n - 1.) <= 1.e-12;
}

C8MI_HD c8mi_env_model const* model_of(EmContext const& c, std::int32_t node) {
  if (node < 0 || static_cast<std::uint32_t>(node) >= c.node_count) return nullptr;
  std::int32_t const m = c.nodes[node].model;
  if (m < 0 || static_cast<std::uint32_t>(m) >= c.model_count) return nullptr;
  return &c.models[m];
}

C8MI_HD bool valid_environment(EmContext const& c) {
  return c.node_count > 0 && c.model_count > 0 && isfinite(c.maximum_magnetic_deflection) &&
         c.maximum_magnetic_deflection > 0. &&
         c.maximum_magnetic_deflection < 0.5 * 3.141592653589793;
}

C8MI_HD bool contains(c8mi_env_node const& s, double const p[3]) {
  if (s.shape_kind == C8MI_SHAPE_UNIVERSE) return true;
  double const d[3] = {p[0] - s.center_m[0], p[1] - s.center_m[1], p[2] - s.center_m[2]};
  return dot3(d, d) < s.radius_m * s.radius_m;
}

C8MI_HD std::int32_t locate(EmContext const& c, double const p[3]) {
  std::int32_t current = 0;
  for (std::uint32_t hop = 0; hop < c.node_count; ++hop) {
    if (current < 0 || static_cast<std::uint32_t>(current) >= c.node_count) return -2;
    c8mi_env_node const& n = c.nodes[current];
    if (!contains(n, p)) return -1;
    std::int32_t next = -1;
    for (std::uint32_t k = 0; k < n.child_count; ++k) {
      std::int32_t const child = c.children[n.child_begin + k];
      if (contains(c.nodes[child], p)) {
        next = child;
        break;
      }
    }
    if (next == -1) return current;
    current = next;
  }
  return -2;
}

struct Intersections {
  double entry, exit;
  bool has;
};

C8MI_HD Intersections intersect(c8mi_env_node const& s, double const p[3],
                                double const v[3]) {
  if (s.shape_kind == C8MI_SHAPE_UNIVERSE) return {-__builtin_inf(), __builtin_inf(), true};
  double const d[3] = {p[0] - s.center_m[0], p[1] - s.center_m[1], p[2] - s.center_m[2]};
  double const vv = dot3(v, v), vd = dot3(v, d);
  double const disc = vd * vd - vv * (dot3(d, d) - s.radius_m * s.radius_m);
  if (!(disc > 0.)) return {0., 0., false};
  double const root = sqrt(disc);
  return {(-vd - root) / vv, (-vd + root) / vv, true};
}

// next_environment_boundary: own exit or nearest child entry.
C8MI_HD void next_boundary(EmContext const& c, double const p[3], double const v[3],
                  std::int32_t current, double limit, double& parameter,
                  std::int32_t& node) {
  double selected = limit;
  std::int32_t selected_node = current;
  c8mi_env_node const& n = c.nodes[current];
  Intersections const own = intersect(n, p, v);
  if (own.has && own.exit <= selected) {
    selected = own.exit;
    selected_node = n.parent;
  }
  for (std::uint32_t k = 0; k < n.child_count; ++k) {
    std::int32_t const child = c.children[n.child_begin + k];
    Intersections const x = intersect(c.nodes[child], p, v);
    if (!x.has) continue;
    if (x.exit > 0. && x.entry <= selected && -x.entry < x.exit) {
      selected = x.entry;
      selected_node = child;
    }
  }
  if (selected < 0.) selected = 0.;
  parameter = selected;
  node = selected_node;
}

C8MI_HD double safe_expm1(double v) { return v == 0. ? v : expm1(v); }
C8MI_HD double safe_log1p(double v) { return v == 0. ? v : log1p(v); }

C8MI_HD double density(c8mi_env_model const& m, double const p[3]) {
  if (m.density_kind == C8MI_DENSITY_HOMOGENEOUS) return m.reference_density_g_per_cm3;
  double const d[3] = {p[0] - m.anchor_m[0], p[1] - m.anchor_m[1], p[2] - m.anchor_m[2]};
  double const h = m.density_kind == C8MI_DENSITY_SLIDING_PLANAR_EXPONENTIAL ? sqrt(dot3(d, d))
                                                                             : dot3(d, m.axis);
  return m.reference_density_g_per_cm3 * exp(m.inverse_scale_length_per_m * (h - m.reference_height_m));
}

C8MI_HD double projected_axis(c8mi_env_model const& m, double const p[3],
                              double const dir[3]) {
  if (m.density_kind == C8MI_DENSITY_FLAT_EXPONENTIAL) return dot3(dir, m.axis);
  double const d[3] = {p[0] - m.anchor_m[0], p[1] - m.anchor_m[1], p[2] - m.anchor_m[2]};
  return dot3(d, dir) / sqrt(dot3(d, d));
}

C8MI_HD double model_grammage(c8mi_env_model const& m, double const p[3],
                              double const dir[3], double length) {
  if (length == 0.) return 0.;
  double const rho = density(m, p);
  if (m.density_kind == C8MI_DENSITY_HOMOGENEOUS) return length * rho * 100.;
  double const proj = projected_axis(m, p, dir);
  if (proj == 0.) return length * rho * 100.;
  return rho * (m.scale_length_m / proj) * safe_expm1(proj * length * m.inverse_scale_length_per_m) * 100.;
}

C8MI_HD double model_arclength(c8mi_env_model const& m, double const p[3],
                               double const dir[3], double grammage) {
  double const rho = density(m, p);
  if (m.density_kind == C8MI_DENSITY_HOMOGENEOUS) return grammage / (rho * 100.);
  double const proj = projected_axis(m, p, dir);
  if (proj == 0.) return grammage / (rho * 100.);
  double const arg = grammage * .01 * m.inverse_scale_length_per_m * proj / rho;
  return arg > -1. ? m.scale_length_m / proj * safe_log1p(arg) : __builtin_inf();
}

struct Layer {
  int status;
  std::int32_t node;
  double density;
};

C8MI_HD Layer query(EmContext const& c, double const p[3], double const* dir,
           std::int32_t logical) {
  if (!valid_environment(c)) return {kInvalid, -1, 0.};
  for (int a = 0; a < 3; ++a)
    if (!isfinite(p[a]) || (dir && !isfinite(dir[a]))) return {kNonFinite, -1, 0.};
  if (dir && !normalized(dir)) return {kNonFinite, -1, 0.};
  std::int32_t const node = logical == -3 ? locate(c, p) : logical;
  c8mi_env_model const* m = model_of(c, node);
  if (!m) return {kOutside, node, 0.};
  double const rho = density(*m, p);
  return {isfinite(rho) && rho > 0. ? kSuccess : kInvalid, node, rho};
}

struct Boundary {
  int status;
  double distance;
  std::int32_t next_node;
};

C8MI_HD Boundary distance_to_boundary(EmContext const& c, double const p[3],
                             double const dir[3], std::int32_t logical) {
  Layer const l = query(c, p, dir, logical);
  if (l.status != kSuccess) return {l.status, 0., -3};
  double param;
  std::int32_t node;
  next_boundary(c, p, dir, l.node, __builtin_inf(), param, node);
  if (!isfinite(param)) return {kNoForward, 0., node};
  return {kSuccess, param, node};
}

struct Grammage {
  int status;
  double value;
};

C8MI_HD Grammage grammage(EmContext const& c, std::int32_t node, double const p[3],
                 double const dir[3], double distance) {
  c8mi_env_model const* m = model_of(c, node);
  if (!valid_environment(c) || !m || !isfinite(distance) || distance < 0. || !normalized(dir))
    return {kNonFinite, 0.};
  double const x = model_grammage(*m, p, dir, distance);
  return {isfinite(x) && x >= 0. ? kSuccess : kGrammageOutOfRange, x};
}

C8MI_HD Grammage arclength(EmContext const& c, std::int32_t node, double const p[3],
                  double const dir[3], double x) {
  c8mi_env_model const* m = model_of(c, node);
  if (!valid_environment(c) || !m || !isfinite(x) || x < 0. || !normalized(dir))
    return {kNonFinite, 0.};
  double const l = model_arclength(*m, p, dir, x);
  return {isfinite(l) && l >= 0. ? kSuccess : kGrammageOutOfRange, l};
}

C8MI_HD Grammage competition_grammage(EmContext const& c, std::int32_t node,
                             double const p[3], double const dir[3],
                             double distance) {
  Grammage r = grammage(c, node, p, dir, distance);
  if (r.status == kGrammageOutOfRange && r.value == __builtin_huge_val() && isfinite(distance) &&
      distance > 0.) {
    c8mi_env_model const* m = model_of(c, node);
    if (m && (m->density_kind == C8MI_DENSITY_SLIDING_PLANAR_EXPONENTIAL ||
              m->density_kind == C8MI_DENSITY_FLAT_EXPONENTIAL)) {
      double const rho = density(*m, p);
      double const growth = projected_axis(*m, p, dir) * m->inverse_scale_length_per_m;
      if (isfinite(rho) && rho > 0. && isfinite(growth) && growth > 0. &&
          isfinite(m->scale_length_m) && m->scale_length_m * m->inverse_scale_length_per_m > 0.)
        r.status = kSuccess;
    }
  }
  return r;
}

// Straight intersection with the observation plane; -1 when there is none.
C8MI_HD double observation_distance(EmContext const& c, double const p[3],
                           double const dir[3]) {
  if (!c.observation_enabled) return -1.;
  double ns = 0.;
  for (int a = 0; a < 3; ++a) {
    if (!isfinite(c.observation_point[a]) || !isfinite(c.observation_normal[a])) return -1.;
    ns += c.observation_normal[a] * c.observation_normal[a];
  }
  if (!isfinite(ns) || fabs(ns - 1.) > 1.e-12) return -1.;
  double disp[3];
  for (int a = 0; a < 3; ++a) {
    if (!isfinite(p[a]) || !isfinite(dir[a])) return -1.;
    disp[a] = c.observation_point[a] - p[a];
  }
  double const den = dot3(c.observation_normal, dir);
  if (den == 0.) return -1.;
  double const d = dot3(c.observation_normal, disp) / den;
  if (!isfinite(d) || !(d > 0.)) return -1.;
  return d;
}

}  // namespace c8mi::em::env
