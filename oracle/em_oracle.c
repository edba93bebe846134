/*
 * c8mi oracle — TEST INFRASTRUCTURE ONLY (see c8_oracle.h).
 *
 * Plain-C FP64 restatement of the EM transport step of CORSIKA8-kokkos and of
 * a serial wavefront driver over it:
 *   environment   corsika/detail/kokkos/environment/{EnvironmentQuery,ModelEvaluation,
 *                 ModelGrammage,StraightIntersections,EnvironmentSearch,ObservationPlane,
 *                 EnvironmentLeptonBoundary,UniformMagneticField}.inl,
 *                 corsika/detail/framework/geometry/NextEnvironmentBoundaryBody.inl
 *   selection     corsika/detail/kokkos/em/algorithms/InteractionSelection.inl:12-215,
 *                 corsika/detail/kokkos/em/tables/{DiscretePhysicsQueries,
 *                 ContinuousPhysicsQueries,TableQueryResult}.inl
 *   photon        PhotonTransportStep.inl:57-261, PhotonFinalStateStep.inl:61-825,
 *                 models/PhotonPairLpm.inl:58-157, models/PhotonPairKinematics.inl:29-72
 *   lepton        LeptonContinuousStep.inl:43-159, LeptonTransportStep.inl:27-513
 *                 (zero field), LeptonVertexSelection.inl:41-199,
 *                 LeptonFinalStateStep.inl:131-1036, models/BremsLpm.inl:58-164
 *   thinning      EmThinning.inl:28-104
 *   fronts        corsika/detail/kokkos/em/execution/{PhotonFront,LeptonFront}.inl
 *                 (stable source order, history IDs from the child prefix sum)
 *   profile       corsika/detail/kokkos/output/algorithms/ProfileAccumulation.inl,
 *                 ProfileProjection.inl, LongitudinalCrossings.inl
 * Built without floating-point contraction (oracle/Makefile).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "c8_oracle.h"
#include "c8mi/c8mi_em.h"
#include "native_internal.h"

/* ---- constants (ScalarPhysicalConstants.hpp:19-22, TransportMass.hpp:26-27,
 *      PhotonPairKinematics.hpp:15, ParticleId.hpp:24, ContinuousPhysicsQueries.hpp:29,
 *      LeptonContinuousStep.hpp:18, LeptonFinalStateStep.hpp:79-81) ---- */
#define C_LIGHT 299792458.
#define ELEMENTARY_CHARGE 1.6021766208e-19
#define MAGNETIC_RIGIDITY 0.2997925146834389
#define ELECTRON_MASS_GEV 0.0005109989461
#define TRANSPORT_ELECTRON_MASS_GEV 5.109989e-4
#define TRANSPORT_MUON_MASS_GEV 1.056584e-1
#define PARTICLE_CUT_TIME_S 10.e-3
#define CONTINUOUS_CUT_SAFETY 0.9999
#define MAX_CONTINUOUS_RELATIVE_LOSS 0.10
#define PHOTON_TWO_PI 6.283185307179586476925286766559
#define PHOTON_PI 3.1415926535897932384626433832795
#define LEPTON_PI 3.141592653589793238462643383279502884
#define LEPTON_TWO_PI 6.283185307179586476925286766559005768
#define PROPOSAL_HALF_PRECISION 1.e-5

/* random domains (RandomDomains.hpp:16-31, PhotonFinalStateRandomDomains.hpp:16-27,
 * LeptonFinalStateRandomDomains.hpp:15-25, EmThinning.hpp:18-19) */
#define RP_DISTANCE 0x454d0001U
#define RP_COLUMN 0x454d0002U
#define RP_SELECTION 0x454d0004U
#define RP_SCATTERING 0x454d0003U
#define D_PAIR_SPLIT 1ULL
#define D_PAIR_AZIMUTH 2ULL
#define D_PAIR_E_POLAR 3ULL
#define D_PAIR_P_POLAR 4ULL
#define D_PAIR_LPM 5ULL
#define D_MU_PAIR_SPLIT (1ULL << 32)              /* PhotonFinalStateRandomDomains.hpp:25-27 */
#define D_MU_PAIR_AZIMUTH ((1ULL << 32) + 1ULL)
#define NATIVE_MUON_PAIR_CUT_FLAG 1U             /* NativeMuonPairCut.hpp:17 */
#define D_COMPTON_AZIMUTH 1ULL
#define D_BREMS_AZIMUTH 1ULL
#define D_BREMS_LPM 2ULL
#define D_ANNIHILATION_RHO 1ULL
#define D_ANNIHILATION_AZIMUTH 2ULL
#define D_IONIZATION_AZIMUTH 1ULL
#define D_THIN_FIRST 0x100ULL
#define D_THIN_SECOND 0x101ULL

enum { P_BREMS = 1000000002, P_IONIZATION = 1000000003, P_EPAIR = 1000000004,
       P_COMPTON = 1000000010, P_ANNIHILATION = 1000000012, P_PHOTON_PAIR = 1000000013,
       P_PHOTON_MU_PAIR = 1000000015, P_PHOTOELECTRIC = 1000000016,
       P_PHOTON_PAIR_FS = 2000000013, P_PHOTON_MU_PAIR_FS = 2000000015 };

/* InteractionStatus (InteractionRecord.hpp:17-45) */
enum { IS_SELECTED = 0, IS_NO_DISCRETE = 1, IS_RESELECT = 2, IS_DISTANCE_SAMPLED = 3,
       IS_PARTICLE_CUT = 4 };
/* TableLookupStatus (TableQueryResult.hpp:18-33) */
enum { TS_SUCCESS = 0, TS_PARTICLE_NOT_FOUND = 1, TS_COLUMN_NOT_FOUND = 2,
       TS_RATE_ENERGY_OOR = 3, TS_INVERSE_CDF = 4, TS_LOSS_ENERGY_OOR = 5,
       TS_LOSS_QUANTILE_OOR = 6, TS_NON_FINITE = 7, TS_INVALID_KIND = 8, TS_INVALID_VIEW = 9,
       TS_CONT_PARTICLE_NOT_FOUND = 10, TS_CONT_ENERGY_OOR = 11, TS_CONT_RANGE_OOR = 12,
       TS_TRANSPORT_CUT = 13 };
/* FallbackReason (FallbackRequest.hpp:22-65) */
enum { FR_UNSUPPORTED_PARTICLE = 0, FR_UNSUPPORTED_GEOMETRY = 2, FR_PARTICLE_TABLE_MISSING = 3,
       FR_COMPONENT_MISSING = 4, FR_RATE_ENERGY_OOR = 5, FR_INVERSE_CDF = 6,
       FR_LOSS_ENERGY_OOR = 7, FR_LOSS_QUANTILE_OOR = 8, FR_INVALID_TABLE_QUERY = 9,
       FR_ZERO_TOTAL_RATE = 11, FR_CPU_ONLY = 12, FR_NOT_IMPLEMENTED = 13,
       FR_INVALID_FINAL_STATE = 14, FR_LPM_UNAVAILABLE = 15, FR_INVALID_DENSITY = 16,
       FR_CONT_ENERGY_OOR = 17, FR_CONT_RANGE_OOR = 18, FR_TRANSPORT_CUT = 19,
       FR_MAGNETIC_TRANSPORT = 22, FR_MAGNETIC_BOUNDARY = 23, FR_GRAMMAGE = 24,
       FR_MOLIERE_UNAVAILABLE = 20, FR_MOLIERE_FAILED = 21,
       FR_INVERSE_GRAMMAGE = 25, FR_VERTEX_LOOKUP = 26, FR_SELECTION_REPLAY = 28 };
/* EnvironmentStatus (EnvironmentQuery.hpp:24-32) */
enum { ES_SUCCESS = 0, ES_INVALID = 1, ES_NON_FINITE = 2, ES_OUTSIDE = 3, ES_NO_FORWARD = 5,
       ES_GRAMMAGE_OOR = 6 };
/* Photon / lepton transport limits (PhotonTransportRecord.hpp:16-23,
 * LeptonTransportRecord.hpp:16-26) */
enum { PL_INTERACTION = 0, PL_LAYER = 1, PL_OBSERVATION = 2, PL_ESCAPED = 3, PL_CUT = 4 };
enum { LL_INTERACTION = 0, LL_CONTINUOUS = 1, LL_CUT = 2, LL_LAYER = 3, LL_OBSERVATION = 4,
       LL_ESCAPED = 5, LL_MAGNETIC = 6, LL_DECAY = 7 };
/* fallback stages */
enum { STAGE_SELECTION = 1, STAGE_TRANSPORT = 2, STAGE_VERTEX = 3, STAGE_FINAL = 4 };

typedef c8mi_particle_state ps_t;
typedef c8mi_native_table_set view_t;

static int finite(double v) { return isfinite(v); }
static double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static int sgn(double v) { return (0. < v) - (v < 0.); }

/* ========================== environment ========================== */

typedef struct {
  const c8mi_environment* e;
} env_t;

static const c8mi_env_model* get_model(const c8mi_environment* e, int32_t node) {
  if (node < 0 || (uint32_t)node >= e->node_count) return NULL;
  int32_t m = e->nodes[node].model;
  if (m < 0 || (uint32_t)m >= e->model_count) return NULL;
  return &e->models[m];
}

/* valid_environment (EnvironmentQuery.inl:53-65) */
static int valid_environment(const c8mi_environment* e) {
  return e->node_count > 0 && e->nodes && e->model_count > 0 && e->models &&
         finite(e->maximum_magnetic_deflection_rad) && e->maximum_magnetic_deflection_rad > 0. &&
         e->maximum_magnetic_deflection_rad < 0.5 * 3.141592653589793;
}

static int normalized(const double d[3]) {
  double n = dot3(d, d);
  return finite(n) && fabs(n - 1.) <= 1.e-12;
}

/* contains (EnvironmentSearch.inl:48-59), spheres and the universe */
static int contains(const c8mi_env_node* s, const double p[3]) {
  if (s->shape_kind == C8MI_SHAPE_UNIVERSE) return 1;
  double d[3] = {p[0] - s->center_m[0], p[1] - s->center_m[1], p[2] - s->center_m[2]};
  return dot3(d, d) < s->radius_m * s->radius_m;
}

/* locate_node (EnvironmentSearch.inl:62-100) from the root */
static int32_t locate_node(const c8mi_environment* e, const double p[3]) {
  int32_t current = 0;
  for (uint32_t hop = 0; hop < e->node_count; ++hop) {
    if (current < 0 || (uint32_t)current >= e->node_count) return -2;
    if (!contains(&e->nodes[current], p)) return -1;
    int32_t next = -1;
    const c8mi_env_node* n = &e->nodes[current];
    for (uint32_t c = 0; c < n->child_count; ++c) {
      int32_t child = e->children[n->child_begin + c];
      if (contains(&e->nodes[child], p)) {
        next = child;
        break;
      }
    }
    if (next == -1) return current;
    current = next;
  }
  return -2;
}

/* straight_intersect (StraightIntersections.inl:72-116), spheres and the universe */
typedef struct {
  double entry, exit;
  int has;
} isect_t;

static isect_t straight_intersect(const c8mi_env_node* s, const double p[3], const double v[3]) {
  isect_t r = {0., 0., 0};
  if (s->shape_kind == C8MI_SHAPE_UNIVERSE) {
    r.entry = -INFINITY;
    r.exit = INFINITY;
    r.has = 1;
    return r;
  }
  double d[3] = {p[0] - s->center_m[0], p[1] - s->center_m[1], p[2] - s->center_m[2]};
  double vv = dot3(v, v), vd = dot3(v, d);
  double disc = vd * vd - vv * (dot3(d, d) - s->radius_m * s->radius_m);
  if (!(disc > 0.)) return r;
  double root = sqrt(disc);
  r.entry = (-vd - root) / vv;
  r.exit = (-vd + root) / vv;
  r.has = 1;
  return r;
}

/* next_environment_boundary (NextEnvironmentBoundaryBody.inl) */
static void next_boundary(const c8mi_environment* e, const double p[3], const double v[3],
                          int32_t current, double limit, double* parameter, int32_t* node) {
  double selected = limit;
  int32_t selected_node = current;
  const c8mi_env_node* n = &e->nodes[current];
  isect_t own = straight_intersect(n, p, v);
  if (own.has && own.exit <= selected) {
    selected = own.exit;
    selected_node = n->parent;
  }
  for (uint32_t c = 0; c < n->child_count; ++c) {
    int32_t child = e->children[n->child_begin + c];
    isect_t x = straight_intersect(&e->nodes[child], p, v);
    if (!x.has) continue;
    if (x.exit > 0. && x.entry <= selected && -x.entry < x.exit) {
      selected = x.entry;
      selected_node = child;
    }
  }
  if (selected < 0.) selected = 0.;
  *parameter = selected;
  *node = selected_node;
}

/* density (ModelEvaluation.inl:61-77) */
static double density(const c8mi_env_model* m, const double p[3]) {
  if (m->density_kind == C8MI_DENSITY_HOMOGENEOUS) return m->reference_density_g_per_cm3;
  double d[3] = {p[0] - m->anchor_m[0], p[1] - m->anchor_m[1], p[2] - m->anchor_m[2]};
  double h = m->density_kind == C8MI_DENSITY_SLIDING_PLANAR_EXPONENTIAL ? sqrt(dot3(d, d))
                                                                         : dot3(d, m->axis);
  return m->reference_density_g_per_cm3 *
         exp(m->inverse_scale_length_per_m * (h - m->reference_height_m));
}

/* projected_axis (ModelGrammage.inl:16-26) */
static double projected_axis(const c8mi_env_model* m, const double p[3], const double dir[3]) {
  if (m->density_kind == C8MI_DENSITY_FLAT_EXPONENTIAL) return dot3(dir, m->axis);
  double d[3] = {p[0] - m->anchor_m[0], p[1] - m->anchor_m[1], p[2] - m->anchor_m[2]};
  return dot3(d, dir) / sqrt(dot3(d, d));
}

/* grammage / arclength closed forms (ModelGrammage.inl:39-84) */
static double model_grammage(const c8mi_env_model* m, const double p[3], const double dir[3],
                             double length) {
  if (length == 0.) return 0.;
  double rho = density(m, p);
  if (m->density_kind == C8MI_DENSITY_HOMOGENEOUS) return length * rho * 100.;
  double proj = projected_axis(m, p, dir);
  if (proj == 0.) return length * rho * 100.;
  return rho * (m->scale_length_m / proj) * expm1(proj * length * m->inverse_scale_length_per_m) *
         100.;
}

static double model_arclength(const c8mi_env_model* m, const double p[3], const double dir[3],
                              double grammage) {
  double rho = density(m, p);
  if (m->density_kind == C8MI_DENSITY_HOMOGENEOUS) return grammage / (rho * 100.);
  double proj = projected_axis(m, p, dir);
  if (proj == 0.) return grammage / (rho * 100.);
  double arg = grammage * .01 * m->inverse_scale_length_per_m * proj / rho;
  return arg > -1. ? m->scale_length_m / proj * log1p(arg) : INFINITY;
}

typedef struct {
  int status;
  int32_t node;
  double density;
} layer_t;

/* query_environment (EnvironmentQuery.inl:73-102) */
static layer_t query_environment(const c8mi_environment* e, const double p[3], const double* dir,
                                 int32_t logical) {
  layer_t r = {ES_INVALID, -1, 0.};
  if (!valid_environment(e)) return r;
  for (int a = 0; a < 3; ++a)
    if (!finite(p[a]) || (dir && !finite(dir[a]))) {
      r.status = ES_NON_FINITE;
      return r;
    }
  if (dir && !normalized(dir)) {
    r.status = ES_NON_FINITE;
    return r;
  }
  int32_t node = logical == -3 ? locate_node(e, p) : logical;
  const c8mi_env_model* m = get_model(e, node);
  r.node = node;
  if (!m) {
    r.status = ES_OUTSIDE;
    return r;
  }
  double rho = density(m, p);
  r.density = rho;
  r.status = finite(rho) && rho > 0. ? ES_SUCCESS : ES_INVALID;
  return r;
}

typedef struct {
  int status;
  double distance;
  int32_t next_node;
} boundary_t;

/* distance_to_environment_boundary (EnvironmentQuery.inl:104-123) */
static boundary_t distance_to_boundary(const c8mi_environment* e, const double p[3],
                                       const double dir[3], int32_t logical) {
  boundary_t r = {ES_INVALID, 0., -3};
  layer_t l = query_environment(e, p, dir, logical);
  if (l.status != ES_SUCCESS) {
    r.status = l.status;
    return r;
  }
  double param;
  int32_t node;
  next_boundary(e, p, dir, l.node, INFINITY, &param, &node);
  r.next_node = node;
  if (!finite(param)) {
    r.status = ES_NO_FORWARD;
    return r;
  }
  r.status = ES_SUCCESS;
  r.distance = param;
  return r;
}

typedef struct {
  int status;
  double value;
} gq_t;

/* environment_grammage / environment_arclength / geometry_competition_grammage
 * (EnvironmentQuery.inl:125-186) */
static gq_t environment_grammage(const c8mi_environment* e, int32_t node, const double p[3],
                                 const double dir[3], double distance) {
  gq_t r = {ES_NON_FINITE, 0.};
  const c8mi_env_model* m = get_model(e, node);
  if (!valid_environment(e) || !m || !finite(distance) || distance < 0. || !normalized(dir))
    return r;
  double x = model_grammage(m, p, dir, distance);
  r.status = finite(x) && x >= 0. ? ES_SUCCESS : ES_GRAMMAGE_OOR;
  r.value = x;
  return r;
}

static gq_t environment_arclength(const c8mi_environment* e, int32_t node, const double p[3],
                                  const double dir[3], double grammage) {
  gq_t r = {ES_NON_FINITE, 0.};
  const c8mi_env_model* m = get_model(e, node);
  if (!valid_environment(e) || !m || !finite(grammage) || grammage < 0. || !normalized(dir))
    return r;
  double l = model_arclength(m, p, dir, grammage);
  r.status = finite(l) && l >= 0. ? ES_SUCCESS : ES_GRAMMAGE_OOR;
  r.value = l;
  return r;
}

static gq_t competition_grammage(const c8mi_environment* e, int32_t node, const double p[3],
                                 const double dir[3], double distance) {
  gq_t r = environment_grammage(e, node, p, dir, distance);
  if (r.status == ES_GRAMMAGE_OOR && r.value == HUGE_VAL && finite(distance) && distance > 0.) {
    const c8mi_env_model* m = get_model(e, node);
    if (m && (m->density_kind == C8MI_DENSITY_SLIDING_PLANAR_EXPONENTIAL ||
              m->density_kind == C8MI_DENSITY_FLAT_EXPONENTIAL)) {
      double rho = density(m, p);
      double growth = projected_axis(m, p, dir) * m->inverse_scale_length_per_m;
      if (finite(rho) && rho > 0. && finite(growth) && growth > 0. && finite(m->scale_length_m) &&
          m->scale_length_m * m->inverse_scale_length_per_m > 0.)
        r.status = ES_SUCCESS;
    }
  }
  return r;
}

/* intersect_observation_plane_straight (ObservationPlane.inl:61-90); returns the
 * distance or -1 when there is no forward intersection. */
static double observation_distance(const c8mi_environment* e, const double p[3],
                                   const double dir[3]) {
  if (!e->observation_enabled) return -1.;
  double ns = 0.;
  for (int a = 0; a < 3; ++a) {
    if (!finite(e->observation_point_m[a]) || !finite(e->observation_normal[a])) return -1.;
    ns += e->observation_normal[a] * e->observation_normal[a];
  }
  if (!finite(ns) || fabs(ns - 1.) > 1.e-12) return -1.;
  double disp[3];
  for (int a = 0; a < 3; ++a) {
    if (!finite(p[a]) || !finite(dir[a])) return -1.;
    disp[a] = e->observation_point_m[a] - p[a];
  }
  double den = dot3(e->observation_normal, dir);
  if (den == 0.) return -1.;
  double d = dot3(e->observation_normal, disp) / den;
  if (!finite(d) || !(d > 0.)) return -1.;
  return d;
}

/* ====================== uniform magnetic field ====================== */
/* maximum_uniform_magnetic_step / advance_uniform_magnetic_field
 * (UniformMagneticField.inl:101-195), lepton_boundary /
 * lepton_observation_distance (EnvironmentLeptonBoundary.inl:16-80) over
 * CurvedBoundaryAccess::intersectShape (CurvedBoundaryAccess.inl:38-126), and
 * the root finders (PolynomialRoots.inl:98-134, StableQuartic.inl) in the
 * double-double arithmetic of the reference's device build
 * (ExtendedPrecision.inl). */

enum { MAG_SUCCESS = 0, MAG_LINEAR = 1, MAG_INVALID = 2, MAG_NONFINITE = 3 };

typedef struct {
  double hi, lo;
} dd_t;

static dd_t dd(double v) {
  dd_t r = {v, 0.};
  return r;
}
static double dd_double(dd_t a) { return a.hi + a.lo; }
static dd_t dd_renorm(double a, double b) {
  double s = a + b;
  dd_t r = {s, b - (s - a)};
  return r;
}
static dd_t dd_add(dd_t a, dd_t b) {
  double s = a.hi + b.hi;
  double bv = s - a.hi;
  double e = (a.hi - (s - bv)) + (b.hi - bv);
  return dd_renorm(s, e + (a.lo + b.lo));
}
static dd_t dd_neg(dd_t a) {
  dd_t r = {-a.hi, -a.lo};
  return r;
}
static dd_t dd_sub(dd_t a, dd_t b) { return dd_add(a, dd_neg(b)); }
static dd_t dd_mul(dd_t a, dd_t b) {
  double p = a.hi * b.hi;
  double e = fma(a.hi, b.hi, -p);
  e = e + (a.hi * b.lo + a.lo * b.hi);
  e = e + a.lo * b.lo;
  return dd_renorm(p, e);
}
static dd_t dd_div(dd_t a, dd_t b) {
  double q = a.hi / b.hi;
  if (!finite(q)) return dd(q);
  dd_t r = dd_sub(a, dd_mul(b, dd(q)));
  double c = r.hi / b.hi;
  r = dd_sub(r, dd_mul(b, dd(c)));
  return dd_add(dd_add(dd(q), dd(c)), dd(r.hi / b.hi));
}
static int dd_lt(dd_t a, dd_t b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }
static int dd_eq(dd_t a, dd_t b) { return a.hi == b.hi && a.lo == b.lo; }
static dd_t dd_abs(dd_t a) { return dd_lt(a, dd(0.)) ? dd_neg(a) : a; }
static dd_t dd_sqrt(dd_t v) {
  double a = sqrt(v.hi);
  if (a == 0. || !finite(a)) return dd(a);
  return dd_add(dd(a), dd_div(dd_sub(v, dd_mul(dd(a), dd(a))), dd(2. * a)));
}
static dd_t dd_cbrt(dd_t v) {
  dd_t a = dd(cbrt(v.hi));
  if (a.hi == 0. || !finite(a.hi)) return a;
  for (int it = 0; it < 2; ++it)
    a = dd_add(a, dd_div(dd_sub(v, dd_mul(dd_mul(a, a), a)), dd_mul(dd_mul(dd(3.), a), a)));
  return a;
}
static dd_t dd_acos(dd_t v) {
  double angle = acos(v.hi);
  if (v.lo == 0.) return dd(angle);
  if (v.hi == 1.) return dd_sqrt(dd_mul(dd(2.), dd_sub(dd(1.), v)));
  if (v.hi == -1.) return dd_sub(dd(3.14159265358979323846), dd_sqrt(dd_mul(dd(2.), dd_add(dd(1.), v))));
  return dd_sub(dd(angle), dd(v.lo / sqrt((1. - v.hi) * (1. + v.hi))));
}
static dd_t dd_cos(dd_t v) { return dd_sub(dd(cos(v.hi)), dd(sin(v.hi) * v.lo)); }
static dd_t dd_fma(dd_t a, dd_t b, dd_t c) { return dd_add(dd_mul(a, b), c); }

typedef struct {
  double v[4];
  unsigned n;
} roots_t;

static void roots_append(roots_t* r, double x) { r->v[r->n++] = x; }
static void roots_append_distinct(roots_t* r, double x) {
  for (unsigned i = 0; i < r->n; ++i)
    if (fabs(x - r->v[i]) <= 8 * 2.220446049250313e-16 * fmax(fabs(x), fabs(r->v[i]))) return;
  if (r->n < 4) roots_append(r, x);
}

/* quadratic_roots (PolynomialRoots.inl:104-134), epsilon 1e-12 */
static roots_t quadratic_roots_dd(dd_t a, dd_t b, dd_t c) {
  const double epsilon = 1.e-12;
  roots_t r = {{0., 0., 0., 0.}, 0u};
  if (dd_lt(dd_abs(a), dd(epsilon))) {
    double bb = dd_double(b), cc = dd_double(c);
    if (bb != 0.) roots_append(&r, -cc / bb);
    return r;
  }
  if (dd_lt(dd_abs(c), dd(epsilon))) {
    double aa = dd_double(a), bb = dd_double(b);
    if (aa != 0.) roots_append(&r, -bb / aa);
    roots_append(&r, 0.);
    return r;
  }
  dd_t a4 = dd_mul(a, dd(4.));
  dd_t multiplied = dd_mul(a4, c);
  dd_t error = dd_fma(a4, c, dd_neg(multiplied));
  dd_t corrected = dd_fma(b, b, dd_neg(multiplied));
  dd_t radicand = dd_add(corrected, error);
  if (dd_lt(dd_abs(radicand), dd(epsilon))) {
    roots_append(&r, dd_double(dd_div(dd_neg(b), dd_mul(dd(2.), a))));
    return r;
  }
  if (dd_lt(radicand, dd(0.))) return r;
  dd_t den = dd_lt(dd(0.), b) ? dd_sub(dd_neg(b), dd_sqrt(radicand)) : dd_add(dd_neg(b), dd_sqrt(radicand));
  dd_t root = dd_div(dd_mul(c, dd(2.)), den);
  roots_append(&r, dd_double(root));
  roots_append(&r, dd_double(dd_div(c, dd_mul(a, root))));
  return r;
}

/* stable_monic_quadratic (StableQuartic.inl) */
static roots_t monic_quadratic_dd(dd_t b, dd_t c) {
  roots_t r = {{0., 0., 0., 0.}, 0u};
  if (dd_eq(c, dd(0.))) {
    roots_append(&r, 0.);
    if (!dd_eq(b, dd(0.))) roots_append(&r, dd_double(dd_neg(b)));
    return r;
  }
  dd_t multiplied = dd_mul(dd(4.), c);
  dd_t disc = dd_fma(b, b, dd_neg(multiplied));
  if (dd_lt(disc, dd(0.))) return r;
  if (dd_eq(disc, dd(0.))) {
    roots_append(&r, dd_double(dd_div(dd_neg(b), dd(2.))));
    return r;
  }
  dd_t s = !dd_lt(b, dd(0.)) ? dd_sqrt(disc) : dd_neg(dd_sqrt(disc));
  dd_t large = dd_div(dd_neg(dd_add(b, s)), dd(2.));
  roots_append(&r, dd_double(large));
  roots_append(&r, dd_double(dd_div(c, large)));
  return r;
}

static dd_t halley_dd(dd_t root, dd_t p, dd_t quadratic, dd_t nc) {
  for (int it = 0; it < 4; ++it) {
    dd_t value = dd_sub(dd_mul(dd_add(dd_mul(dd_add(root, p), root), quadratic), root), nc);
    dd_t derivative = dd_add(dd_mul(dd_add(dd_mul(dd(3.), root), dd_mul(dd(2.), p)), root), quadratic);
    dd_t second = dd_add(dd_mul(dd(6.), root), dd_mul(dd(2.), p));
    dd_t num = dd_mul(dd_mul(dd(2.), value), derivative);
    dd_t den = dd_sub(dd_mul(dd_mul(dd(2.), derivative), derivative), dd_mul(value, second));
    root = dd_sub(root, dd_div(num, den));
  }
  return root;
}

/* stable_resolvent (StableQuartic.inl) */
static dd_t resolvent_dd(dd_t p, dd_t q, dd_t r) {
  dd_t quadratic = dd_sub(dd_div(dd_mul(p, p), dd(4.)), r);
  dd_t nc = dd_div(dd_mul(q, q), dd(8.));
  dd_t zero = dd(0.);
  if (dd_lt(p, zero) && dd_lt(r, zero) && dd_lt(nc, dd_div(dd_mul(dd_neg(p), quadratic), dd(64.))) &&
      dd_lt(nc, dd_div(dd_mul(p, r), dd(8.))))
    return halley_dd(dd_div(nc, quadratic), p, quadratic, nc);
  if (!dd_lt(p, zero) && !dd_lt(quadratic, zero)) {
    dd_t root = dd_cbrt(nc);
    if (dd_lt(zero, p)) {
      dd_t candidate = dd_sqrt(dd_div(nc, p));
      if (dd_lt(candidate, root)) root = candidate;
    }
    if (dd_lt(zero, quadratic)) {
      dd_t candidate = dd_div(nc, quadratic);
      if (dd_lt(candidate, root)) root = candidate;
    }
    return halley_dd(root, p, quadratic, nc);
  }
  dd_t dq = dd_div(dd_sub(dd_mul(dd(3.), quadratic), dd_mul(p, p)), dd(9.));
  dd_t dc = dd_div(dd_add(dd_mul(p, dd_sub(dd_mul(dd(9.), quadratic), dd_mul(dd_mul(dd(2.), p), p))),
                          dd_mul(dd(27.), nc)),
                   dd(54.));
  dd_t disc = dd_fma(dc, dc, dd_mul(dd_mul(dq, dq), dq));
  if (!dd_lt(disc, zero)) {
    dd_t s = !dd_lt(dc, zero) ? dd_sqrt(disc) : dd_neg(dd_sqrt(disc));
    dd_t cube = dd_cbrt(dd_add(dc, s));
    dd_t t = dd_eq(cube, zero) ? zero : dd_sub(cube, dd_div(dq, cube));
    return dd_sub(t, dd_div(p, dd(3.)));
  }
  dd_t argument = dd_div(dc, dd_sqrt(dd_mul(dd_mul(dd_neg(dq), dq), dq)));
  if (dd_lt(dd(1.), argument)) argument = dd(1.);
  if (dd_lt(argument, dd(-1.))) argument = dd(-1.);
  dd_t t = dd_mul(dd_mul(dd(2.), dd_sqrt(dd_neg(dq))), dd_cos(dd_div(dd_acos(argument), dd(3.))));
  return dd_sub(t, dd_div(p, dd(3.)));
}

/* stable_quartic_roots (StableQuartic.inl): x^4 + p x^2 + q x + r */
static roots_t quartic_roots(double p, double q, double r) {
  roots_t out = {{0., 0., 0., 0.}, 0u};
  int pe = -4096, qe = -4096, re = -4096, e;
  if (p != 0.) {
    frexp(p, &e);
    pe = e >= 0 ? (e + 1) / 2 : e / 2;
  }
  if (q != 0.) {
    frexp(q, &e);
    qe = e >= 0 ? (e + 2) / 3 : e / 3;
  }
  if (r != 0.) {
    frexp(r, &e);
    re = e >= 0 ? (e + 3) / 4 : e / 4;
  }
  int exponent = pe > qe ? pe : qe;
  exponent = exponent > re ? exponent : re;
  if (exponent == -4096) {
    roots_append(&out, 0.);
    return out;
  }
  dd_t sp = dd(ldexp(p, -2 * exponent)), sq = dd(ldexp(q, -3 * exponent)), sr = dd(ldexp(r, -4 * exponent));
  double scale = ldexp(1., exponent);
  if (q == 0.) {
    roots_t squares = monic_quadratic_dd(sp, sr);
    for (unsigned i = 0; i < squares.n; ++i) {
      if (squares.v[i] < 0.) continue;
      if (squares.v[i] == 0.) {
        roots_append_distinct(&out, 0.);
      } else {
        double root = dd_double(dd_sqrt(dd(squares.v[i]))) * scale;
        roots_append_distinct(&out, -root);
        roots_append_distinct(&out, root);
      }
    }
  } else {
    dd_t y = resolvent_dd(sp, sq, sr);
    if (!dd_lt(dd(0.), y)) {
      roots_append(&out, INFINITY);
      return out;
    }
    dd_t linear = dd_sqrt(dd_mul(dd(2.), y));
    dd_t common = dd_add(dd_div(sp, dd(2.)), y);
    dd_t difference = dd_div(sq, dd_mul(dd(2.), linear));
    dd_t first, second;
    if (!dd_lt(dd_abs(dd_add(common, difference)), dd_abs(dd_sub(common, difference)))) {
      second = dd_add(common, difference);
      first = dd_div(sr, second);
    } else {
      first = dd_sub(common, difference);
      second = dd_div(sr, first);
    }
    roots_t a = monic_quadratic_dd(linear, first);
    roots_t b = monic_quadratic_dd(dd_neg(linear), second);
    for (unsigned i = 0; i < a.n; ++i) roots_append_distinct(&out, a.v[i] * scale);
    for (unsigned i = 0; i < b.n; ++i) roots_append_distinct(&out, b.v[i] * scale);
  }
  for (unsigned i = 1; i < out.n; ++i) {
    double v = out.v[i];
    unsigned j = i;
    while (j && v < out.v[j - 1]) {
      out.v[j] = out.v[j - 1];
      --j;
    }
    out.v[j] = v;
  }
  return out;
}

typedef struct {
  int status;
  double distance_m, gyroradius_m;
} mlimit_t;

static int magnetic_valid_particle(const ps_t* p, double mass_gev, double charge) {
  if (!finite(p->energy_gev) || !(p->energy_gev > mass_gev) || !finite(mass_gev) || !(mass_gev > 0.) ||
      !finite(charge))
    return 0;
  for (int a = 0; a < 3; ++a)
    if (!finite(p->position_m[a]) || !finite(p->direction[a])) return 0;
  double n2 = dot3(p->direction, p->direction);
  return finite(n2) && fabs(n2 - 1.) <= 1.e-12;
}

static int magnetic_valid_field(const double b[3]) { return finite(b[0]) && finite(b[1]) && finite(b[2]); }

static mlimit_t magnetic_step_limit(const ps_t* p, double mass_gev, double charge, const double b[3],
                                    double max_deflection) {
  mlimit_t out = {MAG_INVALID, 0., 0.};
  if (!magnetic_valid_particle(p, mass_gev, charge) || !magnetic_valid_field(b) || !finite(max_deflection) ||
      !(max_deflection > 0.) || !(max_deflection < 0.5 * 3.141592653589793))
    return out;
  double b2 = dot3(b, b);
  out.status = MAG_LINEAR;
  out.distance_m = INFINITY;
  out.gyroradius_m = INFINITY;
  if (charge == 0. || b2 == 0.) return out;
  double momentum = sqrt((p->energy_gev - mass_gev) * (p->energy_gev + mass_gev));
  double pv[3] = {momentum * p->direction[0], momentum * p->direction[1], momentum * p->direction[2]};
  double along = dot3(pv, b) / b2;
  double perp[3] = {pv[0] - along * b[0], pv[1] - along * b[1], pv[2] - along * b[2]};
  double p_perp = sqrt(dot3(perp, perp));
  if (p_perp < 1.e-9) return out;
  double radius = p_perp / (MAGNETIC_RIGIDITY * fabs(charge) * sqrt(b2));
  out.gyroradius_m = radius;
  if (!finite(radius) || radius > 1.e9) return out;
  double distance = 2. * cos(max_deflection) * sin(max_deflection) * radius;
  if (!finite(distance) || !(distance > 0.)) {
    out.status = MAG_NONFINITE;
    out.distance_m = 0.;
    return out;
  }
  out.status = MAG_SUCCESS;
  out.distance_m = distance;
  return out;
}

static int magnetic_advance(const ps_t* start, double mass_gev, double charge, const double b[3],
                            double distance, ps_t* end) {
  if (!magnetic_valid_particle(start, mass_gev, charge) || !magnetic_valid_field(b) || !finite(distance) ||
      distance < 0.)
    return MAG_INVALID;
  *end = *start;
  if (distance == 0.) return MAG_SUCCESS;
  double momentum = sqrt((start->energy_gev - mass_gev) * (start->energy_gev + mass_gev));
  double coefficient = charge * MAGNETIC_RIGIDITY * distance / momentum;
  const double* u = start->direction;
  double uxb[3] = {u[1] * b[2] - u[2] * b[1], u[2] * b[0] - u[0] * b[2], u[0] * b[1] - u[1] * b[0]};
  double kicked[3], disp[3];
  for (int a = 0; a < 3; ++a) {
    double kick = coefficient * uxb[a];
    kicked[a] = u[a] + kick;
    disp[a] = distance * (u[a] + 0.5 * kick);
    end->position_m[a] = start->position_m[a] + disp[a];
  }
  double kn = sqrt(dot3(kicked, kicked));
  double chord = sqrt(dot3(disp, disp));
  double beta = momentum / start->energy_gev;
  if (!finite(kn) || !(kn > 0.) || !finite(chord) || !finite(beta) || !(beta > 0.)) return MAG_NONFINITE;
  for (int a = 0; a < 3; ++a) end->direction[a] = kicked[a] / kn;
  end->time_s = start->time_s + distance / (beta * C_LIGHT);
  return dot3(uxb, uxb) == 0. || charge == 0. ? MAG_LINEAR : MAG_SUCCESS;
}

typedef struct {
  double position[3], velocity[3], direction[3], field[3];
  double speed, momentum_si, charge_si, gyroradius;
  int linear, incomplete;
} curved_t;

static curved_t make_curved(const ps_t* p, double mass_gev, double charge, const double b[3],
                            const mlimit_t* limit, int linear) {
  curved_t a;
  double momentum = sqrt((p->energy_gev - mass_gev) * (p->energy_gev + mass_gev));
  a.speed = C_LIGHT * momentum / p->energy_gev;
  a.charge_si = charge * ELEMENTARY_CHARGE;
  a.momentum_si = momentum * ELEMENTARY_CHARGE / MAGNETIC_RIGIDITY;
  for (int k = 0; k < 3; ++k) {
    a.position[k] = p->position_m[k];
    a.velocity[k] = a.speed * p->direction[k];
    a.direction[k] = p->direction[k];
    a.field[k] = b[k];
  }
  a.gyroradius = limit->gyroradius_m;
  a.linear = linear;
  a.incomplete = 0;
  return a;
}

/* intersectShape for spheres and the universe (a node of the tree) */
static isect_t curved_sphere(curved_t* a, const c8mi_env_node* s) {
  isect_t none = {0., 0., 0};
  if (a->linear) return straight_intersect(s, a->position, a->velocity);
  if (s->shape_kind == C8MI_SHAPE_UNIVERSE) return none;
  const double* u = a->direction;
  const double* b = a->field;
  double projected[3] = {u[1] * b[2] - u[2] * b[1], u[2] * b[0] - u[0] * b[2], u[0] * b[1] - u[1] * b[0]};
  double delta[3] = {a->position[0] - s->center_m[0], a->position[1] - s->center_m[1],
                     a->position[2] - s->center_m[2]};
  double projected2 = dot3(projected, projected);
  if (projected2 == 0.) return straight_intersect(s, a->position, a->velocity);
  double delta_norm = sqrt(dot3(delta, delta));
  double distance = fabs(delta_norm - s->radius_m);
  if (distance > a->gyroradius) return none;
  if (a->gyroradius > 1000 * distance) return straight_intersect(s, a->position, a->velocity);
  double k = a->charge_si / a->momentum_si;
  double denominator = 4. / (projected2 * k * k);
  double quadratic = (dot3(projected, delta) * k + 1) * denominator;
  double linear = dot3(u, delta) * 2 * denominator;
  double constant = (delta_norm + s->radius_m) * (delta_norm - s->radius_m) * denominator;
  roots_t roots = quartic_roots(quadratic, linear, constant);
  int inside = contains(s, a->position);
  double entry = 0., exit = 0.;
  unsigned count = 0, entries = 0, exits = 0;
  for (unsigned i = 0; i < roots.n; ++i) {
    double d = roots.v[i];
    if (!finite(d)) {
      a->incomplete = 1;
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
  if ((inside && (!entries || !exits)) || (!inside && count < 2)) {
    a->incomplete = 1;
    return none;
  }
  isect_t r = {entry / a->speed, exit / a->speed, 1};
  return r;
}

/* lepton_boundary: 1 on success (distance, next node), 0 on a failed boundary */
static int magnetic_lepton_boundary(const c8mi_environment* e, const ps_t* p, double mass_gev, double charge,
                                    const double b[3], const mlimit_t* limit, double* distance, int32_t* next) {
  curved_t a = make_curved(p, mass_gev, charge, b, limit, limit->status == MAG_LINEAR);
  int32_t current = p->environment_node;
  double selected = limit->distance_m / a.speed;
  int32_t selected_node = current;
  const c8mi_env_node* n = &e->nodes[current];
  isect_t own = curved_sphere(&a, n);
  if (own.has && own.exit <= selected) {
    selected = own.exit;
    selected_node = n->parent;
  }
  for (uint32_t c = 0; c < n->child_count; ++c) {
    int32_t child = e->children[n->child_begin + c];
    isect_t x = curved_sphere(&a, &e->nodes[child]);
    if (!x.has) continue;
    if (x.exit > 0. && x.entry <= selected && -x.entry < x.exit) {
      selected = x.entry;
      selected_node = child;
    }
  }
  if (selected < 0.) selected = 0.;
  *distance = selected * a.speed;
  *next = selected_node;
  return !a.incomplete;
}

/* lepton_observation_distance: the separation-plane branch of intersectShape
 * on the curved trajectory (the access is never linear here) */
static double magnetic_observation_distance(const c8mi_environment* e, const ps_t* p, double mass_gev,
                                            double charge, const double b[3], const mlimit_t* limit) {
  curved_t a = make_curved(p, mass_gev, charge, b, limit, 0);
  const double* u = a.direction;
  double projected[3] = {u[1] * b[2] - u[2] * b[1], u[2] * b[0] - u[0] * b[2], u[0] * b[1] - u[1] * b[0]};
  double delta[3] = {a.position[0] - e->observation_point_m[0], a.position[1] - e->observation_point_m[1],
                     a.position[2] - e->observation_point_m[2]};
  double denominator = a.charge_si * dot3(e->observation_normal, projected);
  double linear = 2 * a.momentum_si * dot3(u, e->observation_normal);
  double constant = 2 * a.momentum_si * dot3(e->observation_normal, delta);
  roots_t r = quadratic_roots_dd(dd(denominator), dd(linear), dd(constant));
  double first = INFINITY;
  for (unsigned i = 0; i < r.n; ++i)
    if (r.v[i] >= 0. && r.v[i] < first) first = r.v[i];
  double entry = first / a.speed;
  return entry > 0. ? entry * a.speed : INFINITY;
}

/* exported for the tests */
unsigned oracle_quartic_roots(double p, double q, double r, double out[4]) {
  roots_t x = quartic_roots(p, q, r);
  for (unsigned i = 0; i < x.n; ++i) out[i] = x.v[i];
  return x.n;
}
unsigned oracle_quadratic_roots(double a, double b, double c, double out[4]) {
  roots_t x = quadratic_roots_dd(dd(a), dd(b), dd(c));
  for (unsigned i = 0; i < x.n; ++i) out[i] = x.v[i];
  return x.n;
}
int oracle_magnetic_step_limit(const c8mi_particle_state* p, double mass_gev, double charge,
                               const double field[3], double max_deflection, double out[2]) {
  mlimit_t l = magnetic_step_limit(p, mass_gev, charge, field, max_deflection);
  out[0] = l.distance_m;
  out[1] = l.gyroradius_m;
  return l.status;
}
int oracle_magnetic_advance(const c8mi_particle_state* p, double mass_gev, double charge, const double field[3],
                            double distance, c8mi_particle_state* end) {
  return magnetic_advance(p, mass_gev, charge, field, distance, end);
}
/* lepton_boundary + lepton_observation_distance over an environment;
 * out = {boundary distance, next node, observation distance}; returns the
 * limit status, or -1 for a failed boundary. */
int oracle_lepton_boundary(const c8mi_environment* e, const c8mi_particle_state* p, double mass_gev,
                           double charge, const double field[3], double out[3]) {
  mlimit_t l = magnetic_step_limit(p, mass_gev, charge, field, e->maximum_magnetic_deflection_rad);
  if (l.status == MAG_INVALID || l.status == MAG_NONFINITE) return l.status;
  double distance;
  int32_t next;
  if (!magnetic_lepton_boundary(e, p, mass_gev, charge, field, &l, &distance, &next)) return -1;
  out[0] = distance;
  out[1] = (double)next;
  out[2] = magnetic_observation_distance(e, p, mass_gev, charge, field, &l);
  return l.status;
}

/* ========================== table facades ========================== */

static int from_native(uint32_t status, int continuous) {
  switch (status) {
    case C8MI_QUERY_SUCCESS: return TS_SUCCESS;
    case C8MI_QUERY_PARTICLE_NOT_FOUND: return continuous ? TS_CONT_PARTICLE_NOT_FOUND
                                                          : TS_PARTICLE_NOT_FOUND;
    case C8MI_QUERY_COLUMN_NOT_FOUND: return TS_COLUMN_NOT_FOUND;
    case C8MI_QUERY_ENERGY_OUT_OF_RANGE: return continuous ? TS_CONT_ENERGY_OOR
                                                           : TS_RATE_ENERGY_OOR;
    case C8MI_QUERY_QUANTILE_OUT_OF_RANGE: return TS_LOSS_QUANTILE_OOR;
    case C8MI_QUERY_UNSUPPORTED_KINEMATICS:
    case C8MI_QUERY_ROOT_NOT_CONVERGED: return TS_INVERSE_CDF;
    case C8MI_QUERY_NON_FINITE_INPUT: return TS_NON_FINITE;
    default: return TS_INVALID_VIEW;
  }
}

typedef struct {
  int status;
  double value;
} tq_t;

/* proposal_fallback_reason (FallbackRequest.inl:14-44) */
static int fallback_reason(int status) {
  switch (status) {
    case TS_PARTICLE_NOT_FOUND: return FR_PARTICLE_TABLE_MISSING;
    case TS_COLUMN_NOT_FOUND: return FR_COMPONENT_MISSING;
    case TS_RATE_ENERGY_OOR: return FR_RATE_ENERGY_OOR;
    case TS_INVERSE_CDF: return FR_INVERSE_CDF;
    case TS_LOSS_ENERGY_OOR: return FR_LOSS_ENERGY_OOR;
    case TS_LOSS_QUANTILE_OOR: return FR_LOSS_QUANTILE_OOR;
    case TS_CONT_PARTICLE_NOT_FOUND: return FR_PARTICLE_TABLE_MISSING;
    case TS_CONT_ENERGY_OOR: return FR_CONT_ENERGY_OOR;
    case TS_CONT_RANGE_OOR: return FR_CONT_RANGE_OOR;
    case TS_TRANSPORT_CUT: return FR_TRANSPORT_CUT;
    default: return FR_INVALID_TABLE_QUERY;
  }
}

static int requires_selected_loss(int reason) {
  return reason == FR_SELECTION_REPLAY || reason == FR_INVERSE_CDF ||
         reason == FR_LOSS_ENERGY_OOR || reason == FR_LOSS_QUANTILE_OOR;
}

typedef struct {
  const view_t* t;
  double em_cut_mev, muon_cut_mev;
} cont_t;

static double transport_mass_gev(int32_t pid) {
  if (pid == 11 || pid == -11) return TRANSPORT_ELECTRON_MASS_GEV;
  if (pid == 13 || pid == -13) return TRANSPORT_MUON_MASS_GEV;
  return 0.;
}
static double secondary_transport_energy(int32_t pid, double native_total, double native_mass) {
  return (native_total - native_mass) + transport_mass_gev(pid);
}
static double mass_correction(int32_t pid, double native_mass, double weight) {
  return weight * (transport_mass_gev(pid) - native_mass);
}
static int is_ee(int32_t pid) { return pid == 11 || pid == -11; }
static int is_muon(int32_t pid) { return pid == 13 || pid == -13; }
static int is_charged_lepton(int32_t pid) { return is_ee(pid) || is_muon(pid); }
/* is_native_muon_pair_cut (NativeMuonPairCut.inl:12-16) */
static int native_muon_pair_cut(const ps_t* p) {
  return p->reserved == NATIVE_MUON_PAIR_CUT_FLAG && is_muon(p->pid) && p->generation > 0 &&
         p->step_id == 0 && p->energy_gev > 0. && p->energy_gev <= transport_mass_gev(p->pid);
}

/* query_continuous_* (ContinuousPhysicsQueries.inl:42-130) */
static tq_t cq_mass(const cont_t* c, int32_t pid) {
  tq_t r = {TS_CONT_PARTICLE_NOT_FOUND, 0.};
  const c8mi_native_utility_column* u = oracle_nq_find_utility(c->t, pid);
  if (!u) return r;
  if (!finite(u->particle_mass_mev) || !(u->particle_mass_mev > 0.)) {
    r.status = TS_INVALID_VIEW;
    return r;
  }
  r.status = TS_SUCCESS;
  r.value = u->particle_mass_mev;
  return r;
}
static tq_t cq_transport_mass(const cont_t* c, int32_t pid) {
  tq_t m = cq_mass(c, pid);
  if (m.status != TS_SUCCESS) return m;
  double tm = 1000. * transport_mass_gev(pid);
  tq_t r = {TS_CONT_PARTICLE_NOT_FOUND, 0.};
  if (!(tm > 0.)) return r;
  r.status = TS_SUCCESS;
  r.value = tm;
  return r;
}
static tq_t cq_transport_cut(const cont_t* c, int32_t pid) {
  double cut = (pid == 13 || pid == -13) ? c->muon_cut_mev : c->em_cut_mev;
  tq_t r = {TS_INVALID_VIEW, 0.};
  if (!finite(cut) || !(cut > 0.)) return r;
  r.status = TS_SUCCESS;
  r.value = cut;
  return r;
}
static tq_t cq_minimum_energy(const cont_t* c, int32_t pid) {
  tq_t r = {TS_CONT_PARTICLE_NOT_FOUND, 0.};
  const c8mi_native_utility_column* u = oracle_nq_find_utility(c->t, pid);
  if (!u) return r;
  tq_t cut = cq_transport_cut(c, pid);
  if (cut.status != TS_SUCCESS) return cut;
  double minimum = (1000. * transport_mass_gev(pid) + cut.value) * CONTINUOUS_CUT_SAFETY;
  if (!finite(u->lower_energy_limit_mev) || u->lower_energy_limit_mev < u->particle_mass_mev ||
      !finite(minimum) || !(minimum > u->particle_mass_mev) ||
      minimum < u->lower_energy_limit_mev || minimum > u->spline.axis.high) {
    r.status = TS_INVALID_VIEW;
    return r;
  }
  r.status = TS_SUCCESS;
  r.value = minimum;
  return r;
}
static tq_t cq_range(const cont_t* c, int32_t pid, double e) {
  c8mi_native_result n = oracle_nq_range(c->t, pid, e);
  tq_t r = {from_native(n.status, 1), n.status == C8MI_QUERY_SUCCESS ? n.value : 0.};
  return r;
}
static tq_t cq_energy_after_loss(const cont_t* c, int32_t pid, double e0, double x) {
  tq_t r = {TS_NON_FINITE, 0.};
  if (!finite(x) || x < 0.) return r;
  tq_t initial = cq_range(c, pid, e0);
  if (initial.status != TS_SUCCESS) return initial;
  if (x >= initial.value && x != 0.) {
    r.status = TS_TRANSPORT_CUT;
    return r;
  }
  c8mi_native_result n = oracle_nq_energy_after_loss(c->t, pid, e0, x);
  r.status = from_native(n.status, 1);
  r.value = n.status == C8MI_QUERY_SUCCESS ? n.value : 0.;
  return r;
}
static tq_t dq_total_rate(const cont_t* c, int32_t pid, double e) {
  c8mi_native_result n = oracle_nq_total_rate(c->t, pid, e);
  tq_t r = {from_native(n.status, 0), n.status == C8MI_QUERY_SUCCESS ? n.value : 0.};
  return r;
}

/* encode / decode of the normalized pair split (PhotonPairKinematics.inl:29-72) */
static int pk_finite(double v) {
  const double maximum = 1.7976931348623157e308;
  return v == v && v <= maximum && v >= -maximum;
}
static int decode_pair_split(double e_gev, double s, double* rho) {
  if (!pk_finite(e_gev) || !pk_finite(s) || !(e_gev > 2. * ELECTRON_MASS_GEV)) return 0;
  const double tol = 64. * 2.2204460492503131e-16;
  if (s < -tol || s > 1. + tol) return 0;
  double b = s < 0. ? 0. : (s > 1. ? 1. : s);
  double lower = ELECTRON_MASS_GEV / e_gev;
  double width = 1. - 2. * lower;
  *rho = lower + width * b;
  return pk_finite(*rho) && *rho >= lower && *rho <= 1. - lower;
}
static int encode_pair_split(double e_gev, double rho, double* s) {
  if (!pk_finite(e_gev) || !pk_finite(rho) || !(e_gev > 2. * ELECTRON_MASS_GEV)) return 0;
  double lower = ELECTRON_MASS_GEV / e_gev;
  double width = 1. - 2. * lower;
  double upper = 1. - lower;
  const double tol = 64. * 2.2204460492503131e-16;
  if (!(width > 0.) || rho < lower - tol || rho > upper + tol) return 0;
  double b = rho < lower ? lower : (rho > upper ? upper : rho);
  *s = (b - lower) / width;
  return pk_finite(*s) && *s >= 0. && *s <= 1.;
}

/* query_discrete_loss_fraction (DiscretePhysicsQueries.inl:69-96) */
static tq_t dq_loss_fraction(const cont_t* c, int32_t pid, int32_t process, uint64_t comp,
                             double e, double q) {
  tq_t r;
  if (process == P_PHOTON_MU_PAIR_FS && pid == 22) {
    c8mi_native_result n = oracle_nq_loss_fraction(c->t, pid, P_PHOTON_MU_PAIR, comp, e, q, 1);
    r.status = from_native(n.status, 0);
    r.value = n.status == C8MI_QUERY_SUCCESS ? n.value : 0.;
    return r;
  }
  if (process == P_PHOTON_PAIR_FS && pid == 22) {
    c8mi_native_result n = oracle_nq_loss_fraction(c->t, pid, P_PHOTON_PAIR, comp, e, q, 1);
    r.status = from_native(n.status, 0);
    r.value = n.status == C8MI_QUERY_SUCCESS ? n.value : 0.;
    if (r.status != TS_SUCCESS) return r;
    double s = 0.;
    if (!encode_pair_split(e / 1000., r.value, &s)) {
      r.status = TS_INVERSE_CDF;
      r.value = 0.;
      return r;
    }
    r.value = s;
    return r;
  }
  c8mi_native_result n = oracle_nq_loss_fraction(c->t, pid, process, comp, e, q, 0);
  r.status = from_native(n.status, 0);
  r.value = n.status == C8MI_QUERY_SUCCESS ? n.value : 0.;
  return r;
}

/* select_rate_column_by_uniform (DiscretePhysicsQueries.inl:24-67) */
typedef struct {
  int status;
  int32_t process;
  uint64_t component;
  double quantile;
  int selected;
} sel_t;

static sel_t select_column(const cont_t* c, int32_t pid, double e, double u) {
  sel_t r = {TS_NON_FINITE, 0, 0, 0., 0};
  if (!finite(u) || u < 0. || u >= 1.) return r;
  c8mi_native_result tot = oracle_nq_selection_total_rate(c->t, pid, e);
  int ts = from_native(tot.status, 0);
  if (ts != TS_SUCCESS) {
    r.status = ts;
    return r;
  }
  double total = tot.value;
  r.status = TS_SUCCESS;
  if (!(total > 0.)) return r;
  double sampled = u * total;
  uint32_t begin, end;
  oracle_nq_dndx_range(c->t, pid, &begin, &end);
  for (uint32_t pos = begin; pos < end; ++pos) {
    const c8mi_native_dndx_column* col = oracle_nq_selection_dndx(c->t, pos);
    if (!col || col->pdg_id != pid) {
      r.status = TS_INVALID_VIEW;
      return r;
    }
    c8mi_native_result rate = {C8MI_QUERY_SUCCESS, 0, 0, 0, 0.};
    if (!oracle_nq_rate_below_threshold(col, e)) rate = oracle_nq_rate_for_column(c->t, col, e);
    int rs = from_native(rate.status, 0);
    double rv = rate.status == C8MI_QUERY_SUCCESS ? rate.value : 0.;
    if (rs != TS_SUCCESS) {
      r.status = rs;
      r.process = col->process_id;
      r.component = col->component_hash;
      return r;
    }
    sampled -= rv;
    if (sampled < 0.) {
      double quantile = (-sampled) / rv;
      r.process = col->process_id;
      r.component = col->component_hash;
      if (!finite(quantile) || quantile < 0. || quantile > 1.) {
        r.status = TS_INVALID_VIEW;
        return r;
      }
      r.quantile = quantile;
      r.selected = 1;
      return r;
    }
  }
  return r;
}

/* process_capability (ProcessCapabilities.inl:45-96) reduced to the classes used */
enum { CAP_PHOTON_PAIR, CAP_PHOTON_MU_PAIR, CAP_BREMS, CAP_COMPTON, CAP_PHOTOELECTRIC,
       CAP_ANNIHILATION, CAP_IONIZATION, CAP_CPU_ONLY, CAP_NOT_IMPLEMENTED };
static int capability(int32_t pid, int32_t process) {
  if (pid == 22 && process == P_PHOTON_MU_PAIR) return CAP_PHOTON_MU_PAIR;
  if (pid == 22 && process == P_PHOTON_PAIR) return CAP_PHOTON_PAIR;
  if (is_ee(pid) && process == P_BREMS) return CAP_BREMS;
  if (pid == 22 && process == P_COMPTON) return CAP_COMPTON;
  if (pid == 22 && process == P_PHOTOELECTRIC) return CAP_PHOTOELECTRIC;
  if (pid == -11 && process == P_ANNIHILATION) return CAP_ANNIHILATION;
  if (is_charged_lepton(pid) && process == P_IONIZATION) return CAP_IONIZATION;
  if (is_charged_lepton(pid) && process == P_EPAIR) return CAP_CPU_ONLY;
  if (is_muon(pid) && process == P_BREMS) return CAP_BREMS;
  switch (process) {
    case 1000000005: case 1000000006: case 1000000007: case 1000000008: case 1000000011:
    case 1000000014:
      return CAP_CPU_ONLY;
    default: return CAP_NOT_IMPLEMENTED;
  }
}
static int requires_replay(int32_t pid, int32_t process, double quantile) {
  int cap = capability(pid, process);
  if (cap == CAP_CPU_ONLY) return 1;
  if (cap == CAP_PHOTON_MU_PAIR || cap == CAP_NOT_IMPLEMENTED) return 0;
  return quantile < 0. || quantile > 1.; /* NativeSelectionReplayEndpointWidth = 0 */
}

/* ========================== records ========================== */

typedef struct {
  ps_t particle;
  int32_t deferred_reason; /* -1: none */
  int32_t status;
  int32_t process_id;
  uint32_t vertex_reached;
  uint64_t component_hash;
  double total_rate;
  double interaction_grammage;
  double mass_density;
  double particle_mass_gev;
  double energy_fraction;
  double process_uniform;
  double loss_quantile;
} inter_t;

typedef struct {
  int flag;
  int reason;
  ps_t particle;
  int32_t process_id;
} fb_t;

static void set_fallback(fb_t* f, const inter_t* in, int reason) {
  f->flag = 1;
  f->reason = reason;
  f->particle = in->particle;
  f->process_id = in->process_id;
}

static double uniform(uint64_t seed, uint64_t shower, const ps_t* p, uint32_t process,
                      uint64_t draw) {
  return oracle_uniform_open01(seed, shower, p->history_id, p->step_id, process, draw, 0);
}

/* ========================== interaction selection ========================== */

typedef struct {
  const view_t* t;
  cont_t cont;
  const c8mi_environment* env;
  const c8mi_em_config* cfg;
  uint64_t seed, shower;
} ctx_t;

/* select_discrete_interaction (InteractionSelection.inl:12-215); muons are not
 * transported by this path and never reach it. */
static void select_interaction(const ctx_t* x, const ps_t* p, inter_t* in, fb_t* fb) {
  memset(in, 0, sizeof *in);
  in->particle = *p;
  in->deferred_reason = -1;
  double e_mev = p->energy_gev * 1000.;
  if (native_muon_pair_cut(p)) { /* InteractionSelection.inl:28-32 */
    in->status = IS_PARTICLE_CUT;
    in->interaction_grammage = INFINITY;
    return;
  }
  if (p->time_s > PARTICLE_CUT_TIME_S) {
    in->status = IS_PARTICLE_CUT;
    in->interaction_grammage = INFINITY;
    return;
  }
  int charged = is_charged_lepton(p->pid);
  int photon = p->pid == 22;
  if (photon && e_mev < x->cont.em_cut_mev) {
    in->status = IS_PARTICLE_CUT;
    in->interaction_grammage = INFINITY;
    return;
  }
  if (charged) {
    tq_t mass = cq_mass(&x->cont, p->pid);
    tq_t tmass = cq_transport_mass(&x->cont, p->pid);
    if (mass.status == TS_SUCCESS) in->particle_mass_gev = mass.value / 1000.;
    tq_t minimum = cq_minimum_energy(&x->cont, p->pid);
    tq_t cut = cq_transport_cut(&x->cont, p->pid);
    int below = tmass.status == TS_SUCCESS && cut.status == TS_SUCCESS &&
                e_mev - tmass.value < cut.value;
    if (below || (minimum.status == TS_SUCCESS && e_mev <= minimum.value)) {
      in->status = IS_PARTICLE_CUT;
      in->interaction_grammage = INFINITY;
      return;
    }
  }
  tq_t total = dq_total_rate(&x->cont, p->pid, e_mev);
  if (total.status != TS_SUCCESS) {
    set_fallback(fb, in, fallback_reason(total.status));
    return;
  }
  in->total_rate = total.value;
  if (!(total.value > 0.)) {
    in->status = IS_NO_DISCRETE;
    in->interaction_grammage = INFINITY;
    return;
  }
  double ud = uniform(x->seed, x->shower, p, RP_DISTANCE, 0);
  in->interaction_grammage = -log(ud) / total.value;
  in->process_uniform = uniform(x->seed, x->shower, p, RP_COLUMN, 0);
  if (charged) {
    in->status = IS_DISTANCE_SAMPLED;
    return;
  }
  double us = uniform(x->seed, x->shower, p, RP_SELECTION, 0);
  sel_t s = select_column(&x->cont, p->pid, e_mev, us);
  if (s.status != TS_SUCCESS) {
    set_fallback(fb, in, fallback_reason(s.status));
    return;
  }
  if (!s.selected) {
    set_fallback(fb, in, FR_ZERO_TOTAL_RATE);
    return;
  }
  in->process_id = s.process;
  in->component_hash = s.component;
  in->loss_quantile = s.quantile;
  if (requires_replay(p->pid, s.process, s.quantile)) {
    in->deferred_reason = FR_SELECTION_REPLAY;
    return; /* status stays Selected (0) */
  }
  tq_t loss = dq_loss_fraction(&x->cont, p->pid, s.process, s.component, e_mev, s.quantile);
  if (loss.status != TS_SUCCESS) {
    int reason = fallback_reason(loss.status);
    if (requires_selected_loss(reason)) {
      in->deferred_reason = reason;
      in->energy_fraction = loss.value;
      return;
    }
    set_fallback(fb, in, reason);
    return;
  }
  in->status = IS_SELECTED;
  in->energy_fraction = loss.value;
}

/* ========================== profile / ledger ========================== */

typedef struct {
  int64_t* photons;
  int64_t* electrons;
  int64_t* positrons;
  int64_t* energy_loss;
  int64_t counters[OR_EM_COUNTERS];
} prof_t;

static long long round_ll(double v) { return (long long)rint(v); }

static void add_fixed(prof_t* pr, int64_t* address, double contribution, double scale) {
  double scaled = contribution * scale;
  if (!finite(scaled) || fabs(scaled) >= 9.2233720368547748e18) {
    pr->counters[OR_EMC_OVERFLOWS]++;
    return;
  }
  long long inc = round_ll(scaled);
  long long prev = *address;
  if ((inc > 0 && prev > INT64_MAX - inc) || (inc < 0 && prev < INT64_MIN - inc)) {
    pr->counters[OR_EMC_OVERFLOWS]++;
    return;
  }
  *address = prev + inc;
}

/* project_profile_grammage (ProfileProjection.inl:12-29) */
static double project_grammage(const c8mi_em_config* c, const double p[3]) {
  double proj = 0.;
  for (int a = 0; a < 3; ++a) proj += (p[a] - c->axis_start_m[a]) * c->axis_direction[a];
  double fb = proj / c->axis_step_m;
  if (fb < 0.) return c->axis_grammage_g_per_cm2[0];
  size_t lower = (size_t)fb;
  size_t upper = lower + 1;
  if (upper >= c->axis_support_count) return c->axis_grammage_g_per_cm2[c->axis_support_count - 1];
  double f = fb - (double)lower;
  return c->axis_grammage_g_per_cm2[upper] * f + c->axis_grammage_g_per_cm2[lower] * (1. - f);
}

double oracle_em_project_grammage(const c8mi_em_config* c, const double p[3]) { return project_grammage(c, p); }

/* longitudinal_profile_bins (LongitudinalCrossings.inl:12-57) for the Both and
 * Forward modes */
static void crossing_bins(double start, double end, size_t count, uint32_t mode, size_t* b,
                          size_t* e) {
  *b = *e = 0;
  if (mode == C8MI_PROFILE_CROSSING_FORWARD && !(end > start)) return;
  if (count == 0 || start == end || !(start >= -1.7976931348623157e308 && start <= 1.7976931348623157e308) ||
      !(end >= -1.7976931348623157e308 && end <= 1.7976931348623157e308))
    return;
  double low = start < end ? start : end, high = start < end ? end : start;
  double first = floor(low) + 1., stop = floor(high) + 1.;
  double limit = (double)count;
  first = first < 0. ? 0. : first;
  stop = stop > limit ? limit : stop;
  if (!(first < stop)) return;
  *b = (size_t)first;
  *e = (size_t)stop;
}

static void particle_profile(prof_t* pr, const c8mi_em_config* c, int32_t pid, double start,
                             double end, double weight) {
  size_t b, e;
  crossing_bins(start / c->profile_bin_width_g_per_cm2, end / c->profile_bin_width_g_per_cm2,
                c->profile_bins, c->profile_crossing_mode, &b, &e);
  if (b == e) return;
  int64_t* h = pid == 22 ? pr->photons : pid == 11 ? pr->electrons : pid == -11 ? pr->positrons
                                                                                : NULL;
  if (!h) {
    pr->counters[OR_EMC_INVALID]++;
    return;
  }
  for (size_t k = b; k < e; ++k) add_fixed(pr, h + k, weight, c->profile_weight_scale);
}

/* accumulate_energy_profile (ProfileAccumulation.inl:110-155) */
static void energy_profile(prof_t* pr, const c8mi_em_config* c, double start, double end,
                           double deposit) {
  if (!(deposit > 0.) || c->profile_bins == 0) return;
  pr->counters[OR_EMC_DEPOSITED_STEPS]++;
  if (start > end) {
    double t = start;
    start = end;
    end = t;
  }
  double delta = end - start;
  int maximum_bin = (int)(c->profile_bins - 1);
  double w = c->profile_bin_width_g_per_cm2;
  int first = (int)(start / w);
  first = first < 0 ? 0 : first;
  first = first > maximum_bin ? maximum_bin : first;
  if (delta < c->profile_energy_loss_threshold_g_per_cm2) {
    add_fixed(pr, pr->energy_loss + first, deposit, c->profile_energy_scale);
    return;
  }
  int last = (int)(end / w);
  last = last < 0 ? 0 : last;
  last = last > maximum_bin ? maximum_bin : last;
  double dens = deposit / delta;
  if (first == last) {
    add_fixed(pr, pr->energy_loss + first, dens * delta, c->profile_energy_scale);
    return;
  }
  add_fixed(pr, pr->energy_loss + first, dens * (1. + (double)first) * w - dens * start,
            c->profile_energy_scale);
  add_fixed(pr, pr->energy_loss + last, dens * end - dens * (double)last * w,
            c->profile_energy_scale);
  for (int bin = first + 1; bin < last; ++bin)
    add_fixed(pr, pr->energy_loss + bin, dens * w, c->profile_energy_scale);
}

/* accumulate_point_energy_profile (ProfileAccumulation.inl:157-175) */
static void point_energy_profile(prof_t* pr, const c8mi_em_config* c, double x, double deposit) {
  if (!(deposit > 0.) || c->profile_bins == 0) return;
  int maximum_bin = (int)(c->profile_bins - 1);
  int bin = (int)(x / c->profile_bin_width_g_per_cm2);
  bin = bin < 0 ? 0 : bin;
  bin = bin > maximum_bin ? maximum_bin : bin;
  pr->counters[OR_EMC_DEPOSITED_STEPS]++;
  add_fixed(pr, pr->energy_loss + bin, deposit, c->profile_energy_scale);
}

static void ledger_add(prof_t* pr, const c8mi_em_config* c, int term, double value) {
  add_fixed(pr, &pr->counters[term], value, c->profile_energy_scale);
}

/* ========================== photon transport ========================== */

typedef struct {
  inter_t in; /* updated interaction (vertex state when reached) */
  ps_t start, end;
  int limit;
  double cut_deposit;
  int observed_before_cut;
} ptrans_t;

static void advance_straight(const ps_t* s, double d, ps_t* e) {
  *e = *s;
  for (int a = 0; a < 3; ++a) e->position_m[a] = s->position_m[a] + d * s->direction[a];
  e->time_s = s->time_s + d / C_LIGHT;
}

/* transport_photon (PhotonTransportStep.inl:57-261) */
static void transport_photon(const ctx_t* x, const inter_t* in, ptrans_t* out, fb_t* fb) {
  const c8mi_environment* env = x->env;
  ps_t start = in->particle;
  memset(out, 0, sizeof *out);
  if (start.pid != 22) {
    set_fallback(fb, in, FR_UNSUPPORTED_PARTICLE);
    return;
  }
  if (in->status != IS_SELECTED && in->status != IS_NO_DISCRETE && in->status != IS_PARTICLE_CUT) {
    set_fallback(fb, in, FR_INVALID_FINAL_STATE);
    return;
  }
  layer_t layer = query_environment(env, start.position_m, start.direction, start.environment_node);
  if (layer.status != ES_SUCCESS) {
    set_fallback(fb, in, FR_UNSUPPORTED_GEOMETRY);
    return;
  }
  start.environment_node = layer.node;
  if (in->status == IS_PARTICLE_CUT) {
    out->in = *in;
    out->in.vertex_reached = 0;
    out->start = start;
    out->end = start;
    out->end.step_id++;
    out->limit = PL_CUT;
    out->cut_deposit = start.energy_gev;
    return;
  }
  boundary_t bnd = distance_to_boundary(env, start.position_m, start.direction, start.environment_node);
  double obs = observation_distance(env, start.position_m, start.direction);
  int has_b = bnd.status == ES_SUCCESS, has_o = obs > 0.;
  if (!has_b && !has_o) {
    set_fallback(fb, in, FR_UNSUPPORTED_GEOMETRY);
    return;
  }
  int obs_wins = has_o && (!has_b || obs < bnd.distance);
  double bd = obs_wins ? obs : bnd.distance;
  gq_t bg = competition_grammage(env, layer.node, start.position_m, start.direction, bd);
  if (bg.status != ES_SUCCESS) {
    set_fallback(fb, in, FR_GRAMMAGE);
    return;
  }
  out->in = *in;
  out->in.vertex_reached = 0;
  out->start = start;
  int reaches = in->status == IS_SELECTED && finite(in->interaction_grammage) &&
                in->interaction_grammage >= 0. && in->interaction_grammage <= bg.value;
  if (reaches) {
    gq_t dist = environment_arclength(env, layer.node, start.position_m, start.direction,
                                      in->interaction_grammage);
    if (dist.status != ES_SUCCESS || dist.value > bd * (1. + 1.e-12)) {
      set_fallback(fb, in, FR_INVERSE_GRAMMAGE);
      return;
    }
    advance_straight(&start, dist.value, &out->end);
    layer_t v = query_environment(env, out->end.position_m, out->end.direction, start.environment_node);
    if (v.status != ES_SUCCESS || !finite(v.density) || !(v.density > 0.)) {
      set_fallback(fb, in, FR_VERTEX_LOOKUP);
      return;
    }
    out->limit = PL_INTERACTION;
    out->in.particle = out->end;
    out->in.mass_density = v.density;
    if (out->end.time_s > PARTICLE_CUT_TIME_S) {
      out->end.step_id++;
      out->limit = PL_CUT;
      out->cut_deposit = start.energy_gev;
    } else {
      out->in.vertex_reached = 1;
    }
    return;
  }
  if (!finite(bg.value)) {
    set_fallback(fb, in, FR_GRAMMAGE);
    return;
  }
  advance_straight(&start, bd, &out->end);
  out->end.step_id++;
  if (obs_wins) {
    out->limit = PL_OBSERVATION;
  } else {
    int32_t node = bnd.next_node;
    out->end.environment_node = node;
    const c8mi_env_model* m = get_model(env, node);
    if (!m) {
      out->limit = PL_ESCAPED;
    } else {
      layer_t next = query_environment(env, out->end.position_m, out->end.direction, node);
      if (next.status != ES_SUCCESS || node == layer.node) {
        set_fallback(fb, in, FR_UNSUPPORTED_GEOMETRY);
        return;
      }
      out->limit = PL_LAYER;
      out->end.medium_id = m->medium_id;
    }
  }
  if (out->end.time_s > PARTICLE_CUT_TIME_S) {
    out->observed_before_cut = out->limit == PL_OBSERVATION;
    out->limit = PL_CUT;
    out->cut_deposit = start.energy_gev;
  }
}

/* ========================== final-state helpers ========================== */

static double bounded_cosine(double v) {
  double m = v < 1. ? v : 1.;
  return -1. > m ? -1. : m;
}

/* sauter_cosine (PhotonFinalStateStep.inl:61-71) */
static double sauter_cosine(double e, double u) {
  double m2 = ELECTRON_MASS_GEV * ELECTRON_MASS_GEV;
  double t = e * e - m2;
  double p = sqrt(0. > t ? 0. : t);
  double coord = 2. * u - 1.;
  double emp = m2 / (e + p);
  double omc = emp * (1. - coord) / (p * coord + e);
  return bounded_cosine(1. - omc);
}

/* compton_cosines (PhotonFinalStateStep.inl:72-96) */
static int compton_cosines(double e, double v, double* pc, double* ec) {
  double scattered = e * (1. - v);
  double pe2 = 2. * v * e * ELECTRON_MASS_GEV + v * v * e * e;
  if (!finite(e) || !(e > 0.) || !finite(v) || !(v > 0.) || !(v < 1.) || !(scattered > 0.) ||
      !(pe2 > 0.))
    return 0;
  *pc = 1. - (v * ELECTRON_MASS_GEV) / scattered;
  *ec = v * (e + ELECTRON_MASS_GEV) / sqrt(pe2);
  const double tol = 1.e-12;
  if (!finite(*pc) || !finite(*ec) || *pc < -1. - tol || *pc > 1. + tol || *ec < -1. - tol ||
      *ec > 1. + tol)
    return 0;
  *pc = bounded_cosine(*pc);
  *ec = bounded_cosine(*ec);
  return 1;
}

/* photoelectric_energy (PhotonFinalStateStep.inl:97-133) */
static int photoelectric_energy(const c8mi_photon_pair_lpm* s, uint64_t comp, double e,
                                double* total, double* fraction) {
  if (!finite(e) || !(e > 0.) || s->component_count == 0 || s->component_count > 16 ||
      !finite(s->fine_structure_constant) || !(s->fine_structure_constant > 0.))
    return 0;
  double z = 0.;
  for (uint32_t i = 0; i < s->component_count; ++i)
    if (s->components[i].component_hash == comp) {
      z = s->components[i].nuclear_charge;
      break;
    }
  if (!finite(z) || !(z > 0.)) return 0;
  double za = z * s->fine_structure_constant;
  double binding = za * za * ELECTRON_MASS_GEV / 2.;
  double kinetic = e - binding;
  if (!finite(binding) || !finite(kinetic) || !(kinetic >= 0.)) return 0;
  *total = ELECTRON_MASS_GEV + kinetic;
  *fraction = kinetic / e;
  return finite(*total) && *total >= ELECTRON_MASS_GEV && finite(*fraction) && *fraction >= 0. &&
         *fraction <= 1.;
}

/* deflect_photon_child (PhotonFinalStateStep.inl:134-160) and deflect_lepton_child
 * (LeptonFinalStateStep.inl:86-116; the lepton form bounds the cosine first) */
static void deflect(const double in[3], double cosine, double azimuth, double out[3], int bound) {
  double t = in[0] * in[0] + in[1] * in[1];
  double tr = sqrt(0. > t ? 0. : t);
  double cp = 1., sp = 0.;
  if (tr > 0.) {
    cp = in[0] / tr;
    sp = in[1] / tr;
  }
  double ct = in[2], st = tr;
  double rx[3] = {ct * cp, ct * sp, -st};
  double ry[3] = {-sp, cp, 0.};
  double c = cosine;
  if (bound) {
    double m = 1. < c ? 1. : c;
    c = -1. > m ? -1. : m;
  }
  double s2 = (1. - c) * (1. + c);
  double s = sqrt(0. > s2 ? 0. : s2);
  double lx = s * cos(azimuth), ly = s * sin(azimuth);
  double z2 = 1. - lx * lx - ly * ly;
  double lz = sqrt(0. > z2 ? 0. : z2);
  if (c < 0.) lz = -lz;
  for (int a = 0; a < 3; ++a) out[a] = lz * in[a] + lx * rx[a] + ly * ry[a];
}

/* apply_em_thinning (EmThinning.inl:28-104) */
typedef struct {
  int status; /* 0 not applied, 1 Hillas, 2 statistical, 3 invalid */
  uint32_t keep;
  double w1, w2;
} thin_t;

static thin_t em_thinning(const c8mi_em_thinning* c, double pe, double pw, double e1, double e2,
                          double u1, double u2) {
  thin_t r = {3, 0, 0., 0.};
  if (!finite(pe) || !(pe > 0.) || !finite(pw) || pw < 0. || !finite(e1) || !(e1 > 0.) ||
      !finite(e2) || !(e2 > 0.) || !finite(u1) || !(u1 > 0.) || !(u1 < 1.) || !finite(u2) ||
      !(u2 > 0.) || !(u2 < 1.))
    return r;
  thin_t unchanged = {0, 3U, pw, pw};
  if (c->enabled == 0) return unchanged;
  if (!finite(c->threshold_gev) || !(c->threshold_gev > 0.) || !finite(c->maximum_weight) ||
      !(c->maximum_weight > 0.))
    return r;
  if (pw >= c->maximum_weight || pe > c->threshold_gev) return unchanged;
  double sum = e1 + e2;
  if (!finite(sum) || !(sum > 0.)) return r;
  double p1 = e1 / sum, p2 = e2 / sum;
  double f1 = 1. / p1, f2 = 1. / p2, fm = c->maximum_weight / pw;
  thin_t o = {0, 0U, 0., 0.};
  if (f1 <= fm && f2 <= fm) {
    o.status = 1;
    if (u1 <= p1) {
      o.keep = 1U;
      o.w1 = pw * f1;
    } else {
      o.keep = 2U;
      o.w2 = pw * f2;
    }
  } else {
    o.status = 2;
    double l1 = fmin(f1, fm), l2 = fmin(f2, fm);
    if (u1 * l1 <= 1.) {
      o.keep |= 1U;
      o.w1 = pw * l1;
    }
    if (u2 * l2 <= 1.) {
      o.keep |= 2U;
      o.w2 = pw * l2;
    }
  }
  if (c->erase_zero_weight == 0) {
    o.keep = 3U;
  } else {
    if (o.w1 == 0.) o.keep &= ~1U;
    if (o.w2 == 0.) o.keep &= ~2U;
  }
  return o;
}

static int two_child_thinning(const ctx_t* x, const ps_t* parent, int32_t process, double e1,
                              double e2, int lepton, thin_t* t) {
  double u1 = uniform(x->seed, x->shower, parent, (uint32_t)process, D_THIN_FIRST);
  double u2 = uniform(x->seed, x->shower, parent, (uint32_t)process, D_THIN_SECOND);
  c8mi_em_thinning cfg = x->cfg->thinning;
  if (lepton && !is_ee(parent->pid)) cfg.enabled = 0;
  *t = em_thinning(&cfg, parent->energy_gev, parent->weight, e1, e2, u1, u2);
  return t->status != 3;
}

/* photon_pair_lpm_suppression_factor (PhotonPairLpm.inl:58-155); status 0 on success */
static int pair_lpm(const c8mi_photon_pair_lpm* s, uint64_t comp, double e, double xx,
                    double rho, double* prob) {
  if (s->component_count == 0 || s->component_count > 16 ||
      !finite(s->baseline_mass_density_g_per_cm3) || !(s->baseline_mass_density_g_per_cm3 > 0.) ||
      !finite(s->molecular_density_per_cm3) || !(s->molecular_density_per_cm3 > 0.) ||
      !finite(s->sum_charge) || !(s->sum_charge > 0.) || !finite(s->e_lpm_mev) ||
      !(s->e_lpm_mev > 0.) || !finite(s->classical_electron_radius_cm) ||
      !(s->classical_electron_radius_cm > 0.) || !finite(s->fine_structure_constant) ||
      !(s->fine_structure_constant > 0.))
    return 2;
  if (!finite(e) || !(e > 0.) || !finite(xx) || !(xx > 0.) || !(xx < 1.) || !finite(rho) ||
      !(rho > 0.))
    return 3;
  const c8mi_photon_pair_lpm_component* c = NULL;
  for (uint32_t i = 0; i < s->component_count; ++i)
    if (s->components[i].component_hash == comp) {
      c = &s->components[i];
      break;
    }
  if (!c) return 1;
  if (!finite(c->nuclear_charge) || !(c->nuclear_charge > 0.) ||
      !finite(c->radiation_log_constant) || !(c->radiation_log_constant > 0.))
    return 2;
  const double Pi = 3.141592653589793238462643383279502884;
  const double SqrtTwo = 1.414213562373095048801688724209698079;
  double dc = rho / s->baseline_mass_density_g_per_cm3;
  double zm = pow(c->nuclear_charge, -1. / 3.);
  double s1 = 1. / (zm * c->radiation_log_constant);
  s1 = s1 * s1 * SqrtTwo;
  double sp = 0.125 * sqrt(s->e_lpm_mev / (dc * e * xx * (1. - xx)));
  double h = log(sp) / log(s1);
  double xi = 1.;
  if (sp < s1) xi = 2.;
  else if (sp < 1.) xi = 1. + h - 0.08 * (1. - h) * (1. - (1. - h) * (1. - h)) / log(s1);
  double g0 = s->classical_electron_radius_cm / (s->fine_structure_constant * xx);
  double gamma = 1. + 4. * Pi * s->sum_charge * s->classical_electron_radius_cm * g0 * g0 *
                          s->molecular_density_per_cm3 * dc;
  double ss = sp / sqrt(xi) * gamma;
  double s2 = ss * ss;
  double fi, g;
  if (ss < 1.54954)
    fi = 1. - exp(-6. * ss * (1. + (3. - Pi) * ss) + s2 * ss / (0.623 + 0.796 * ss + 0.658 * s2));
  else
    fi = 1. - 0.012 / (s2 * s2);
  if (ss < 0.710390) {
    double psi = 1. - exp(-4. * ss - 8. * s2 / (1. + 3.936 * ss + 4.97 * s2 - 0.05 * s2 * ss +
                                                7.50 * s2 * s2));
    g = 3. * psi - 2. * fi;
  } else if (ss < 0.904912) {
    g = 36. * s2 / (36. * s2 + 1.);
  } else {
    g = 1. - 0.022 / (s2 * s2);
  }
  double p = ((xi / 3.) * (g / (gamma * gamma) + 2. * (xx * xx + (1. - xx) * (1. - xx)) * fi / gamma)) /
             (1. - (4. / 3.) * xx * (1. - xx));
  if (!finite(p) || p < 0.) return 4;
  *prob = p;
  return 0;
}

/* brems_lpm_suppression_factor (BremsLpm.inl:58-162); status 0 on success */
static int brems_lpm(const c8mi_brems_lpm* s, uint64_t comp, double e, double v, double rho,
                     double* prob) {
  if (s->component_count == 0 || s->component_count > 16 ||
      !finite(s->baseline_mass_density_g_per_cm3) || !(s->baseline_mass_density_g_per_cm3 > 0.) ||
      !finite(s->molecular_density_per_cm3) || !(s->molecular_density_per_cm3 > 0.) ||
      !finite(s->sum_charge) || !(s->sum_charge > 0.) || !finite(s->e_lpm_mev) ||
      !(s->e_lpm_mev > 0.) || !finite(s->lepton_mass_mev) || !(s->lepton_mass_mev > 0.) ||
      !finite(s->electron_mass_mev) || !(s->electron_mass_mev > 0.) ||
      !finite(s->muon_mass_mev) || !(s->muon_mass_mev > 0.) ||
      !finite(s->classical_electron_radius_cm) || !(s->classical_electron_radius_cm > 0.) ||
      !finite(s->fine_structure_constant) || !(s->fine_structure_constant > 0.))
    return 2;
  if (!finite(e) || !(e > s->lepton_mass_mev) || !finite(v) || !(v > 0.) || !(v < 1.) ||
      !finite(rho) || !(rho > 0.))
    return 3;
  const c8mi_brems_lpm_component* c = NULL;
  for (uint32_t i = 0; i < s->component_count; ++i)
    if (s->components[i].component_hash == comp) {
      c = &s->components[i];
      break;
    }
  if (!c) return 1;
  if (!finite(c->nuclear_charge) || !(c->nuclear_charge > 0.) || !finite(c->atomic_mass_number) ||
      !(c->atomic_mass_number > 0.) || !finite(c->radiation_log_constant) ||
      !(c->radiation_log_constant > 0.))
    return 2;
  const double Pi = 3.141592653589793238462643383279502884;
  const double SqrtTwo = 1.414213562373095048801688724209698079;
  double dc = rho / s->baseline_mass_density_g_per_cm3;
  double zm = pow(c->nuclear_charge, -1. / 3.);
  double dn = 1.54 * pow(c->atomic_mass_number, 0.27);
  double na = s->lepton_mass_mev / s->muon_mass_mev * dn;
  double s1 = s->electron_mass_mev / (s->lepton_mass_mev * zm * c->radiation_log_constant);
  s1 = s1 * s1 * (1. + na * na) * SqrtTwo;
  double sp = 0.125 * sqrt(s->e_lpm_mev * v / (dc * e * (1. - v)));
  double h = log(sp) / log(s1);
  double xi = 1.;
  if (sp < s1) xi = 2.;
  else if (sp < 1.) xi = 1. + h - 0.08 * (1. - h) * (1. - (1. - h) * (1. - h)) / log(s1);
  double g0 = s->classical_electron_radius_cm * s->electron_mass_mev /
              (s->fine_structure_constant * s->lepton_mass_mev * v);
  double gamma = 1. + 4. * Pi * s->sum_charge * s->classical_electron_radius_cm * g0 * g0 *
                          s->molecular_density_per_cm3 * dc;
  double ss = sp / sqrt(xi) * gamma;
  double s2 = ss * ss;
  double fi, g;
  if (ss < 1.54954)
    fi = 1. - exp(-6. * ss * (1. + (3. - Pi) * ss) + s2 * ss / (0.623 + 0.796 * ss + 0.658 * s2));
  else
    fi = 1. - 0.012 / (s2 * s2);
  if (ss < 0.710390) {
    double psi = 1. - exp(-4. * ss - 8. * s2 / (1. + 3.936 * ss + 4.97 * s2 - 0.05 * s2 * ss +
                                                7.50 * s2 * s2));
    g = 3. * psi - 2. * fi;
  } else if (ss < 0.904912) {
    g = 36. * s2 / (36. * s2 + 1.);
  } else {
    g = 1. - 0.022 / (s2 * s2);
  }
  double p = ((xi / 3.) * (v * v * g / (gamma * gamma) + 2. * (1. + (1. - v) * (1. - v)) * fi / gamma)) /
             ((4. / 3.) * (1. - v) + v * v);
  if (!finite(p) || p < 0.) return 4;
  *prob = p;
  return 0;
}

/* ========================== shower state ========================== */

/* Class of a queued particle, which orders the input of a call: the reference
 * uploads the device pending-queue prefix (particles the other species' fronts
 * appended) ahead of the host particles (WavefrontQueue.inl:207-260), and the
 * host particles are the ones left over by the species' previous call
 * (remainingPhotons_/remainingLeptons_, staged by EmRouter.inl:520-521,570-571)
 * followed by newly staged ones. */
enum { CLS_PENDING = 0, CLS_REMAINING = 1, CLS_HOST = 2 };

typedef struct {
  ps_t* items;
  uint8_t* cls;
  uint64_t count, capacity;
} queue_t;

/* The reference's call schedule for one shower (EmRouter.inl:339-416,
 * PhotonCascade.inl:286-300, LeptonCascade.inl:240-260): a photon call of at
 * most 16 fronts, then a lepton call of at most 1024 fronts, repeated; each
 * call reserves its own history range from the shower's counter (the stack's
 * reserveTransportHistoryIds, EmRouter.inl:627-679) and ends early below the
 * minimum batch (after its first front) or, for leptons, when the front could
 * exhaust the range. */
typedef struct {
  int phase;         /* 0 photon call, 1 lepton call */
  int open;          /* a call of `phase` holds a reservation */
  uint64_t fronts;   /* fronts of the open call */
  uint64_t counter;  /* next history ID to reserve */
  uint64_t next;     /* next secondary history ID of the open call */
  uint64_t limit;    /* exclusive limit of the open lepton call */
  uint64_t min_batch;
  uint64_t photon_call_fronts;  /* fronts per call (16 and 1024 in the reference) */
  uint64_t lepton_call_fronts;
  uint64_t lepton_multiplier;   /* lepton reservation max(3 n, min(chunk, multiplier 3 n)) */
  uint64_t lepton_chunk;
} sched_t;

enum { PHOTON_CALL_FRONTS = 16, LEPTON_CALL_FRONTS = 1024, BUCKET_MINIMUM = 256 };

typedef struct {
  ctx_t x;
  queue_t photons, leptons;
  sched_t sched;
  prof_t prof;
  ps_t* observations;
  uint64_t observation_count;
  c8mi_em_fallback* fallbacks;
  uint64_t fallback_count;
  or_em_stats stats;
  /* optional radio tracks of the e+- steps, [track_count][9] */
  double* tracks;
  uint64_t track_count, track_capacity;
  int record_tracks, track_overflow;
  /* optional vertex log (test probe) */
  or_vertex* vertices;
  uint64_t vertex_count, vertex_capacity;
  int record_vertices, vertex_overflow;
  /* the writers' first-interaction snapshot */
  c8mi_em_first_interaction first;
  uint64_t first_candidates;
} shower_t;

static or_vertex* log_vertex(shower_t* s, int kind, const ps_t* parent) {
  if (!s->record_vertices) return NULL;
  if (s->vertex_count == s->vertex_capacity) {
    uint64_t cap = s->vertex_capacity ? 2 * s->vertex_capacity : 1024;
    or_vertex* grown = realloc(s->vertices, cap * sizeof *grown);
    if (!grown) {
      s->vertex_overflow = 1;
      return NULL;
    }
    s->vertices = grown;
    s->vertex_capacity = cap;
  }
  or_vertex* v = &s->vertices[s->vertex_count++];
  memset(v, 0, sizeof *v);
  v->kind = kind;
  v->pid = parent->pid;
  v->energy = parent->energy_gev;
  for (int a = 0; a < 3; ++a) v->direction[a] = parent->direction[a];
  return v;
}

static void log_child(or_vertex* v, int32_t pid, double energy, const double direction[3]) {
  if (!v || v->count >= 3) return;
  v->child_pid[v->count] = pid;
  v->child_energy[v->count] = energy;
  for (int a = 0; a < 3; ++a) v->child_direction[v->count][a] = direction[a];
  v->count++;
}

static void log_thinning(or_vertex* v, uint32_t status, uint32_t keep, double parent_weight, double w1,
                         double w2) {
  if (!v) return;
  v->thin_status = status;
  v->keep = keep;
  v->parent_weight = parent_weight;
  v->child_weight[0] = w1;
  v->child_weight[1] = w2;
}

/* The first final state of a generation-zero parent (PhotonFront.inl:95-111,
 * LeptonFinalStateStep.inl:609-615): children before thinning, with the
 * parent's weight and history ID; the first candidate of the shower wins. */
static void capture_first(shower_t* s, const ps_t* parent, int32_t process, const ps_t* a, const ps_t* b) {
  if (parent->generation != 0) return;
  if (s->first_candidates++ != 0) return;
  memset(&s->first, 0, sizeof s->first);
  s->first.parent_at_vertex = *parent;
  s->first.secondaries[0] = *a;
  s->first.secondaries[0].weight = parent->weight;
  /* the photoelectric electron is snapshot with its assigned history ID
   * (PhotonFinalStateStep.inl:649-660); other children before their IDs */
  if (process != P_PHOTOELECTRIC) s->first.secondaries[0].history_id = parent->history_id;
  if (b) {
    s->first.secondaries[1] = *b;
    s->first.secondaries[1].weight = parent->weight;
    s->first.secondaries[1].history_id = parent->history_id;
  }
  s->first.process_id = process;
  s->first.secondary_count = b ? 2u : 1u;
}

/* Compact radio track of one e+- step (the make_radio_track_kinematics input,
 * RadioProjection.inl:145-168): start and end positions and times, weight
 * signed by charge (negative for e-). */
static void record_track(shower_t* s, const ps_t* start, const ps_t* end) {
  if (!s->record_tracks) return;
  if (s->track_count == s->track_capacity) {
    uint64_t cap = s->track_capacity ? 2 * s->track_capacity : 4096;
    double* grown = realloc(s->tracks, cap * 9 * sizeof(double));
    if (!grown) {
      s->track_overflow = 1;
      return;
    }
    s->tracks = grown;
    s->track_capacity = cap;
  }
  double* t = s->tracks + 9 * s->track_count++;
  for (int k = 0; k < 3; ++k) {
    t[k] = start->position_m[k];
    t[3 + k] = end->position_m[k];
  }
  t[6] = start->time_s;
  t[7] = end->time_s;
  t[8] = start->pid == 11 ? -start->weight : start->weight;
}

static void push(queue_t* q, const ps_t* p, int cls) {
  if (q->count == q->capacity) {
    q->capacity = q->capacity ? 2 * q->capacity : 1024;
    q->items = (ps_t*)realloc(q->items, q->capacity * sizeof(ps_t));
    q->cls = (uint8_t*)realloc(q->cls, q->capacity);
  }
  q->cls[q->count] = (uint8_t)cls;
  q->items[q->count++] = *p;
}

static void queue_free(queue_t* q) {
  free(q->items);
  free(q->cls);
  memset(q, 0, sizeof *q);
}

static void record_observation(shower_t* s, const ps_t* p) {
  if (s->observation_count < s->x.cfg->observation_capacity && s->observations)
    s->observations[s->observation_count] = *p;
  s->observation_count++;
}

static void record_fallback(shower_t* s, const fb_t* f, uint32_t stage) {
  s->prof.counters[OR_EMC_FALLBACK_REASON0 + (f->reason & 31)]++;
  ledger_add(&s->prof, s->x.cfg, OR_EMC_FALLBACK_ENERGY, f->particle.energy_gev * f->particle.weight);
  if (s->fallback_count < s->x.cfg->fallback_capacity && s->fallbacks) {
    c8mi_em_fallback* o = &s->fallbacks[s->fallback_count];
    memset(o, 0, sizeof *o);
    o->particle = f->particle;
    o->reason = f->reason;
    o->process_id = f->process_id;
    o->stage = stage;
  }
  s->fallback_count++;
  s->stats.fallbacks++;
}

static void terminal_energy(shower_t* s, double e, int observed) {
  if (!(e >= 0.) || !finite(e)) {
    s->prof.counters[OR_EMC_INVALID]++;
    return;
  }
  ledger_add(&s->prof, s->x.cfg, observed ? OR_EMC_OBSERVED : OR_EMC_ESCAPED, e);
}

static void thinning_ledger(shower_t* s, int status, uint32_t multiplicity, uint32_t kept,
                            double parent_energy_weight, double kept_energy_weight) {
  if (status == 1) s->prof.counters[OR_EMC_THIN_HILLAS]++;
  else if (status == 2) s->prof.counters[OR_EMC_THIN_STATISTICAL]++;
  s->prof.counters[OR_EMC_THIN_DISCARDED] += multiplicity - kept;
  ledger_add(&s->prof, s->x.cfg, OR_EMC_THINNING_ENERGY, parent_energy_weight - kept_energy_weight);
}

/* ========================== photon front ========================== */

/* One photon: selection, transport, final state, profile and outputs in the
 * order of the reference's materialisation (PhotonFront.inl:137-397). Outputs:
 * continuing photon (next), charged children and history IDs. */
typedef struct {
  ps_t next[2];
  uint32_t next_count;
  ps_t charged[2];
  uint32_t charged_count;
  uint32_t child_count;
} pout_t;

static void photon_source(shower_t* s, const ps_t* p, pout_t* o, uint64_t child_offset,
                          uint64_t first_history) {
  const ctx_t* x = &s->x;
  const c8mi_em_config* cfg = x->cfg;
  memset(o, 0, sizeof *o);
  inter_t in;
  fb_t fb = {0, 0, {0}, 0};
  select_interaction(x, p, &in, &fb);
  if (fb.flag) {
    record_fallback(s, &fb, STAGE_SELECTION);
    return;
  }
  ptrans_t tr;
  transport_photon(x, &in, &tr, &fb);
  if (fb.flag) {
    record_fallback(s, &fb, STAGE_TRANSPORT);
    return;
  }
  s->stats.photon_steps++;
  /* accumulate_photon_profile_step (ProfileAccumulation.inl:228-257) */
  double w = tr.start.weight;
  double gs = project_grammage(cfg, tr.start.position_m);
  double ge = project_grammage(cfg, tr.end.position_m);
  particle_profile(&s->prof, cfg, tr.start.pid, gs, ge, w);
  point_energy_profile(&s->prof, cfg, ge, tr.cut_deposit * w);
  s->prof.counters[OR_EMC_STEPS]++;
  if (tr.limit == PL_CUT) {
    s->prof.counters[OR_EMC_PHOTON_CUTS]++;
    ledger_add(&s->prof, cfg, OR_EMC_CUT_DEPOSIT, tr.cut_deposit * w);
    if (tr.observed_before_cut) {
      ledger_add(&s->prof, cfg, OR_EMC_OVERLAP, tr.cut_deposit * w);
      terminal_energy(s, tr.end.energy_gev * w, 1);
      record_observation(s, &tr.end);
    }
    return;
  }
  if (tr.limit == PL_OBSERVATION) {
    terminal_energy(s, tr.end.energy_gev * w, 1);
    record_observation(s, &tr.end);
    return;
  }
  if (tr.limit == PL_ESCAPED) {
    terminal_energy(s, tr.end.energy_gev * w, 0);
    record_observation(s, &tr.end);
    return;
  }
  if (tr.limit == PL_LAYER) {
    o->next[o->next_count++] = tr.end;
    return;
  }
  /* interaction vertex: classify_photon_final_state (PhotonFinalStateStep.inl:273-558) */
  s->stats.photon_interactions++;
  const inter_t* vi = &tr.in;
  const ps_t* parent = &vi->particle;
  if (vi->deferred_reason >= 0) {
    fb_t f;
    if (vi->vertex_reached != 1 || !requires_selected_loss(vi->deferred_reason))
      set_fallback(&f, vi, FR_INVALID_FINAL_STATE);
    else
      set_fallback(&f, vi, vi->deferred_reason);
    record_fallback(s, &f, STAGE_FINAL);
    return;
  }
  int cap = capability(parent->pid, vi->process_id);
  fb_t f;
  if (cap != CAP_PHOTON_PAIR && cap != CAP_PHOTON_MU_PAIR && cap != CAP_COMPTON &&
      cap != CAP_PHOTOELECTRIC) {
    set_fallback(&f, vi, cap == CAP_CPU_ONLY ? FR_CPU_ONLY : FR_NOT_IMPLEMENTED);
    record_fallback(s, &f, STAGE_FINAL);
    return;
  }
  if (cap == CAP_PHOTON_MU_PAIR) {
    /* PhotonFinalStateStep.inl:326-382 (classification), 679-731 (children):
     * x = E(mu+)/E from the native split CDF, both muons at polar angle m/E,
     * opposite azimuths, the parent's weight, no thinning */
    double mass = x->cont.t->constants.muon_mass_mev / 1000.;
    if (!finite(parent->energy_gev) || !finite(parent->weight) || parent->weight < 0. || !(mass > 0.) ||
        parent->generation == 0xffffffffU || parent->step_id == 0xffffffffffffffffULL) {
      set_fallback(&f, vi, FR_INVALID_FINAL_STATE);
      record_fallback(s, &f, STAGE_FINAL);
      return;
    }
    double su = uniform(x->seed, x->shower, parent, P_PHOTON_MU_PAIR, D_MU_PAIR_SPLIT);
    tq_t split = dq_loss_fraction(&x->cont, 22, P_PHOTON_MU_PAIR_FS, vi->component_hash,
                                  parent->energy_gev * 1000., su);
    if (split.status != TS_SUCCESS) {
      set_fallback(&f, vi, FR_INVALID_FINAL_STATE);
      record_fallback(s, &f, STAGE_FINAL);
      return;
    }
    double minus_total = (1. - split.value) * parent->energy_gev;
    double plus_total = split.value * parent->energy_gev;
    if (!finite(split.value) || !(minus_total > 0.) || !(plus_total > 0.)) {
      set_fallback(&f, vi, FR_INVALID_FINAL_STATE);
      record_fallback(s, &f, STAGE_FINAL);
      return;
    }
    double az = uniform(x->seed, x->shower, parent, P_PHOTON_MU_PAIR, D_MU_PAIR_AZIMUTH) * PHOTON_TWO_PI;
    double k = (parent->energy_gev * 1000.) / (mass * 1000.);
    double cosine = cos(1. / k);
    if (cosine == -1.) cosine *= -1.;
    ps_t mu = *parent, mp = *parent;
    mu.pid = 13;
    mp.pid = -13;
    mu.energy_gev = secondary_transport_energy(13, minus_total, mass);
    mp.energy_gev = secondary_transport_energy(-13, plus_total, mass);
    mu.parent_history_id = mp.parent_history_id = parent->history_id;
    mu.generation = mp.generation = parent->generation + 1;
    mu.step_id = mp.step_id = 0;
    mu.reserved = mu.energy_gev <= transport_mass_gev(13) ? NATIVE_MUON_PAIR_CUT_FLAG : 0U;
    mp.reserved = mp.energy_gev <= transport_mass_gev(-13) ? NATIVE_MUON_PAIR_CUT_FLAG : 0U;
    mu.weight = mp.weight = parent->weight;
    deflect(parent->direction, cosine, az, mu.direction, 0);
    deflect(parent->direction, cosine, fmod(az + PHOTON_PI, PHOTON_TWO_PI), mp.direction, 0);
    {
      or_vertex* lv = log_vertex(s, 1, parent);
      if (lv) {
        lv->process = vi->process_id;
        lv->target_mass = mass;
        lv->v = split.value;
        log_child(lv, 13, minus_total, mu.direction);
        log_child(lv, -13, plus_total, mp.direction);
      }
      log_thinning(lv, 0U, 3U, parent->weight, parent->weight, parent->weight);
    }
    capture_first(s, parent, P_PHOTON_MU_PAIR, &mu, &mp);
    mu.history_id = first_history + child_offset;
    mp.history_id = first_history + child_offset + 1;
    o->charged[o->charged_count++] = mu;
    o->charged[o->charged_count++] = mp;
    o->child_count = 2;
    s->stats.secondaries += 2;
    ledger_add(&s->prof, cfg, OR_EMC_MASS_CORRECTION, mass_correction(13, mass, 2. * parent->weight));
    thinning_ledger(s, 0, 2, 2, 0., 0.);
    return;
  }
  const c8mi_photon_pair_lpm* lpm = &cfg->photon_pair_lpm;
  if (cap == CAP_COMPTON) {
    double pc, ec;
    if (parent->generation == 0xffffffffU || parent->step_id == 0xffffffffffffffffULL ||
        !compton_cosines(parent->energy_gev, vi->energy_fraction, &pc, &ec)) {
      set_fallback(&f, vi, FR_INVALID_FINAL_STATE);
      record_fallback(s, &f, STAGE_FINAL);
      return;
    }
    double v = vi->energy_fraction;
    double az_u = uniform(x->seed, x->shower, parent, P_COMPTON, D_COMPTON_AZIMUTH);
    thin_t th;
    double e1 = parent->energy_gev * (1. - v);
    double e2 = secondary_transport_energy(11, ELECTRON_MASS_GEV + parent->energy_gev * v,
                                           ELECTRON_MASS_GEV);
    if (!two_child_thinning(x, parent, P_COMPTON, e1, e2, 0, &th)) {
      set_fallback(&f, vi, FR_INVALID_FINAL_STATE);
      record_fallback(s, &f, STAGE_FINAL);
      return;
    }
    double az = az_u * PHOTON_TWO_PI;
    ps_t ph = *parent, el = *parent;
    ph.energy_gev = e1;
    ph.parent_history_id = parent->history_id;
    ph.generation = parent->generation + 1;
    ph.step_id = 0;
    ph.reserved = 0;
    ph.weight = th.w1;
    deflect(parent->direction, pc, az, ph.direction, 0);
    el.pid = 11;
    el.energy_gev = e2;
    el.parent_history_id = parent->history_id;
    el.generation = parent->generation + 1;
    el.step_id = 0;
    el.reserved = 0;
    el.weight = th.w2;
    deflect(parent->direction, ec, fmod(az + PHOTON_PI, PHOTON_TWO_PI), el.direction, 0);
    {
      or_vertex* lv = log_vertex(s, 1, parent);
      if (lv) {
        lv->process = vi->process_id;
        lv->target_mass = ELECTRON_MASS_GEV;
        lv->v = v;
        log_child(lv, 22, e1, ph.direction);
        log_child(lv, 11, ELECTRON_MASS_GEV + parent->energy_gev * v, el.direction);
      }
      log_thinning(lv, (uint32_t)th.status, th.keep, parent->weight, th.w1, th.w2);
    }
    capture_first(s, parent, P_COMPTON, &ph, &el);
    uint32_t k = 0;
    if (th.keep & 1U) {
      ph.history_id = first_history + child_offset + k++;
      o->next[o->next_count++] = ph;
    }
    if (th.keep & 2U) {
      el.history_id = first_history + child_offset + k++;
      o->charged[o->charged_count++] = el;
    }
    o->child_count = k;
    s->stats.secondaries += k;
    double kept_lepton_w = (th.keep & 2U) ? th.w2 : 0.;
    ledger_add(&s->prof, cfg, OR_EMC_MASS_CORRECTION, mass_correction(11, ELECTRON_MASS_GEV, kept_lepton_w));
    ledger_add(&s->prof, cfg, OR_EMC_MEDIUM_REST_MASS, ELECTRON_MASS_GEV * w);
    thinning_ledger(s, th.status, 2, k, (e1 + e2) * w,
                    ((th.keep & 1U) ? e1 * th.w1 : 0.) + ((th.keep & 2U) ? e2 * th.w2 : 0.));
    return;
  }
  if (cap == CAP_PHOTOELECTRIC) {
    double etot, frac;
    if (parent->generation == 0xffffffffU || !finite(vi->energy_fraction) ||
        fabs(vi->energy_fraction - 1.) > 1.e-12 ||
        !photoelectric_energy(lpm, vi->component_hash, parent->energy_gev, &etot, &frac)) {
      set_fallback(&f, vi, FR_INVALID_FINAL_STATE);
      record_fallback(s, &f, STAGE_FINAL);
      return;
    }
    ps_t el = *parent;
    el.pid = 11;
    el.energy_gev = secondary_transport_energy(11, ELECTRON_MASS_GEV + parent->energy_gev * frac,
                                               ELECTRON_MASS_GEV);
    el.parent_history_id = parent->history_id;
    el.history_id = first_history + child_offset;
    el.generation = parent->generation + 1;
    el.step_id = 0;
    el.reserved = 0;
    el.weight = parent->weight;
    capture_first(s, parent, P_PHOTOELECTRIC, &el, NULL);
    o->charged[o->charged_count++] = el;
    o->child_count = 1;
    s->stats.secondaries += 1;
    ledger_add(&s->prof, cfg, OR_EMC_MASS_CORRECTION, mass_correction(11, ELECTRON_MASS_GEV, parent->weight));
    ledger_add(&s->prof, cfg, OR_EMC_MEDIUM_REST_MASS, ELECTRON_MASS_GEV * w);
    ledger_add(&s->prof, cfg, OR_EMC_PHOTOELECTRIC_BINDING, tr.end.energy_gev * (1. - frac) * w);
    thinning_ledger(s, 0, 1, 1, 0., 0.);
    return;
  }
  /* photon pair */
  if (!finite(parent->energy_gev) || parent->energy_gev < 2. * ELECTRON_MASS_GEV ||
      parent->generation == 0xffffffffU || parent->step_id == 0xffffffffffffffffULL) {
    set_fallback(&f, vi, FR_INVALID_FINAL_STATE);
    record_fallback(s, &f, STAGE_FINAL);
    return;
  }
  double su = uniform(x->seed, x->shower, parent, P_PHOTON_PAIR, D_PAIR_SPLIT);
  tq_t split = dq_loss_fraction(&x->cont, parent->pid, P_PHOTON_PAIR_FS, vi->component_hash,
                                parent->energy_gev * 1000., su);
  double rho = 0.;
  if (split.status != TS_SUCCESS) {
    set_fallback(&f, vi, fallback_reason(split.status));
    record_fallback(s, &f, STAGE_FINAL);
    return;
  }
  if (!decode_pair_split(parent->energy_gev, split.value, &rho)) {
    set_fallback(&f, vi, FR_INVALID_FINAL_STATE);
    record_fallback(s, &f, STAGE_FINAL);
    return;
  }
  double ee = parent->energy_gev * rho;
  double pe = parent->energy_gev - ee;
  if (!finite(rho) || !(rho > 0.) || !(rho < 1.) || ee < ELECTRON_MASS_GEV || pe < ELECTRON_MASS_GEV) {
    set_fallback(&f, vi, FR_INVALID_FINAL_STATE);
    record_fallback(s, &f, STAGE_FINAL);
    return;
  }
  double az_u = uniform(x->seed, x->shower, parent, P_PHOTON_PAIR, D_PAIR_AZIMUTH);
  double eu = uniform(x->seed, x->shower, parent, P_PHOTON_PAIR, D_PAIR_E_POLAR);
  double pu = uniform(x->seed, x->shower, parent, P_PHOTON_PAIR, D_PAIR_P_POLAR);
  double survival = 0.;
  int ls = pair_lpm(lpm, vi->component_hash, parent->energy_gev * 1000., rho, vi->mass_density,
                    &survival);
  if (ls != 0) {
    set_fallback(&f, vi, ls == 3 ? FR_INVALID_DENSITY : FR_LPM_UNAVAILABLE);
    record_fallback(s, &f, STAGE_FINAL);
    return;
  }
  double lu = uniform(x->seed, x->shower, parent, P_PHOTON_PAIR, D_PAIR_LPM);
  if (lu > survival) {
    ps_t cont = *parent;
    cont.step_id++;
    o->next[o->next_count++] = cont;
    return;
  }
  thin_t th;
  double e1 = secondary_transport_energy(11, ee, ELECTRON_MASS_GEV);
  double e2 = secondary_transport_energy(-11, pe, ELECTRON_MASS_GEV);
  if (!two_child_thinning(x, parent, P_PHOTON_PAIR, e1, e2, 0, &th)) {
    set_fallback(&f, vi, FR_INVALID_FINAL_STATE);
    record_fallback(s, &f, STAGE_FINAL);
    return;
  }
  double az = az_u * PHOTON_TWO_PI;
  double ec = sauter_cosine(ee, eu), pcos = sauter_cosine(pe, pu);
  ps_t el = *parent, po = *parent;
  el.pid = 11;
  el.energy_gev = e1;
  el.parent_history_id = parent->history_id;
  el.generation = parent->generation + 1;
  el.step_id = 0;
  el.reserved = 0;
  el.weight = th.w1;
  deflect(parent->direction, ec, az, el.direction, 0);
  po.pid = -11;
  po.energy_gev = e2;
  po.parent_history_id = parent->history_id;
  po.generation = parent->generation + 1;
  po.step_id = 0;
  po.reserved = 0;
  po.weight = th.w2;
  deflect(parent->direction, pcos, fmod(az + PHOTON_PI, PHOTON_TWO_PI), po.direction, 0);
  {
    or_vertex* lv = log_vertex(s, 1, parent);
    if (lv) {
      lv->process = vi->process_id;
      lv->v = rho;
      log_child(lv, 11, ee, el.direction);
      log_child(lv, -11, pe, po.direction);
    }
    log_thinning(lv, (uint32_t)th.status, th.keep, parent->weight, th.w1, th.w2);
  }
  capture_first(s, parent, P_PHOTON_PAIR, &el, &po);
  uint32_t k = 0;
  if (th.keep & 1U) {
    el.history_id = first_history + child_offset + k++;
    o->charged[o->charged_count++] = el;
  }
  if (th.keep & 2U) {
    po.history_id = first_history + child_offset + k++;
    o->charged[o->charged_count++] = po;
  }
  o->child_count = k;
  s->stats.secondaries += k;
  double kept = ((th.keep & 1U) ? th.w1 : 0.) + ((th.keep & 2U) ? th.w2 : 0.);
  ledger_add(&s->prof, cfg, OR_EMC_MASS_CORRECTION, mass_correction(11, ELECTRON_MASS_GEV, kept));
  thinning_ledger(s, th.status, 2, k, (e1 + e2) * w,
                  ((th.keep & 1U) ? e1 * th.w1 : 0.) + ((th.keep & 2U) ? e2 * th.w2 : 0.));
}

/* ========================== Molière multiple scattering ==========================
 * sample_moliere_scattering_angle_2d with an empty interpolation view (the
 * analytical correction functions), apply_moliere_direction and its
 * step-compatible form (MoliereScattering.inl:219-688), and the scattering stage
 * of a completed e+- step (MoliereStep.inl:11-101). */

enum { MS_SUCCESS = 0, MS_NO_DEFLECTION = 1, MS_INVALID_SNAPSHOT = 2, MS_INVALID_INPUT = 3,
       MS_NON_CONVERGENT = 4, MS_NON_FINITE = 5 };
#define MOLIERE_PI 3.141592653589793238462643383279502884
#define MOLIERE_SQRT2 1.414213562373095048801688724209698079
#define MOLIERE_MIN_B 4.5
#define MOLIERE_MIN_B_RHS 2.9959226032237258

typedef struct {
  double inverse_scale[C8MI_MOLIERE_MAX_COMPONENTS];
  double inverse_b[C8MI_MOLIERE_MAX_COMPONENTS];
  double density_weight[C8MI_MOLIERE_MAX_COMPONENTS];
  double prefactor;
  double beta_squared;
} mol_state_t;

static int moliere_valid(const c8mi_moliere_snapshot* m) {
#define POS(v) (finite(v) && (v) > 0.)
  if (m->component_count == 0 || m->component_count > C8MI_MOLIERE_MAX_COMPONENTS ||
      m->maximum_weight_index >= m->component_count || !POS(m->particle_mass_mev) ||
      !POS(m->electron_mass_mev) || !POS(m->fine_structure_constant) || !POS(m->avogadro_per_mol) ||
      !POS(m->hbar_mev_s) || !POS(m->speed_of_light_cm_per_s) || !POS(m->z_squared_over_a_average) ||
      !POS(m->inverse_weight_zz_sum) || !finite(m->euler_mascheroni))
    return 0;
  for (uint32_t i = 0; i < m->component_count; ++i) {
    const c8mi_moliere_component* c = &m->components[i];
    if (!POS(c->nuclear_charge) || !POS(c->weight_zz) || !POS(c->chi0_squared_mev2) ||
        !POS(c->coulomb_correction))
      return 0;
  }
#undef POS
  for (int i = 0; i < 70; ++i)
    if (!finite(m->c1[i]) || !finite(m->c2[i])) return 0;
  for (int i = 0; i < 13; ++i)
    if (!finite(m->c2_large[i]) || !finite(m->s2_large[i])) return 0;
  for (int i = 0; i < 15; ++i)
    if (!finite(m->c1_integral_large[i])) return 0;
  return 1;
}

/* Acklam's rational approximation of the normal quantile (the Newton seed). */
double oracle_moliere_normal_ppf(double p) {
  static const double a[] = {-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                             1.383577518672690e+02, -3.066479806614716e+01, 2.506628277459239e+00};
  static const double b[] = {-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                             6.680131188771972e+01, -1.328068155288572e+01};
  static const double c[] = {-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                             -2.549732539343734e+00, 4.374664141464968e+00, 2.938163982698783e+00};
  static const double d[] = {7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                             3.754408661907416e+00};
  const double lo = 0.02425, hi = 1. - lo;
  if (p < lo) {
    double q = sqrt(-2. * log(p));
    return (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
           ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.);
  }
  if (p <= hi) {
    double q = p - 0.5, r = q * q;
    return (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
           (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1.);
  }
  double q = sqrt(-2. * log(1. - p));
  return -(((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
         ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.);
}

/* B - ln B = rhs on the B > 1 branch: fourth-order asymptotic seed, one Halley step. */
double oracle_moliere_solve_b(double rhs) {
  if (!finite(rhs) || rhs < MOLIERE_MIN_B_RHS) return 0.;
  double l = log(rhs), ir = 1. / rhs, ir2 = ir * ir, l2 = l * l, l3 = l2 * l;
  double b = rhs + l + l * ir + l * (2. - l) * 0.5 * ir2 + l * (6. - 9. * l + 2. * l2) / 6. * ir2 * ir +
             l * (12. - 36. * l + 22. * l2 - 3. * l3) / 12. * ir2 * ir2;
  if (!finite(b) || !(b > 1.)) return 0.;
  double f = b - log(b) - rhs, ib = 1. / b, df = 1. - ib;
  double den = 2. * df * df - f * ib * ib;
  if (!finite(den) || den == 0.) return 0.;
  b -= 2. * f * df / den;
  return finite(b) && b >= MOLIERE_MIN_B ? b : 0.;
}

double oracle_moliere_f1(const c8mi_moliere_snapshot* m, double x) {
  if (x > 12.) return 0.5 * sqrt(MOLIERE_PI) / (pow(x, 1.5) * pow(1. - 4.5 / x, 2. / 3.));
  double sum = m->c1[69];
  for (int i = 68; i >= 0; --i) sum = sum * x + m->c1[i];
  return sum;
}

double oracle_moliere_f2(const c8mi_moliere_snapshot* m, double x) {
  if (x > 4.25 * 4.25) {
    if (x <= 6.5 * 6.5)
      return 0.00013567765224589459194769192063035 * x - 0.0022635502525409950842771866774683 * sqrt(x) +
             0.0098037758070269476889935233998585;
    double sum = 0.;
    for (int i = 2; i < 13; ++i) sum += m->c2_large[i] * (0.5 * log(x) + m->s2_large[i]) * pow(x, -(i + 0.5));
    return sum;
  }
  double sum = m->c2[69];
  for (int i = 68; i >= 0; --i) sum = sum * x + m->c2[i];
  return sum;
}

double oracle_moliere_integral_f1(const c8mi_moliere_snapshot* m, double x) {
  if (x > 12.) {
    double sum = m->c1_integral_large[14];
    for (int i = 13; i >= 0; --i) sum = m->c1_integral_large[i] + sum / x;
    return sum;
  }
  double sum = m->c1[69] / (2. * 69. + 1.);
  for (int i = 68; i >= 0; --i) sum = sum * x + m->c1[i] / (2. * i + 1.);
  return sum * sqrt(x > 0. ? x : 0.);
}

double oracle_moliere_integral_f2(const c8mi_moliere_snapshot* m, double x) {
  if (x > 4.25 * 4.25) {
    if (x <= 6.5 * 6.5)
      return -0.00026360133958801203364619158975302 * x + 0.0039965027465457608410459577896745 * sqrt(x) -
             0.016305842044996649714549974419242;
    double sum = 0.;
    for (int i = 2; i < 13; ++i)
      sum += -0.5 * m->c2_large[i] / i * (0.5 / i + 0.5 * log(x) + m->s2_large[i]) * pow(x, -i);
    return sum;
  }
  double sum = m->c2[69] / (2. * 69. + 1.);
  for (int i = 68; i >= 0; --i) sum = sum * x + m->c2[i] / (2. * i + 1.);
  return sum * sqrt(x > 0. ? x : 0.);
}

/* ---- interpolation view (MoliereScattering.inl:118-216, 350-385) ---- */

#define MOLIERE_TABLE_UPPER 20.

/* evaluate_interpolation: cubic of the interval holding x on [0, 20]. */
static double moliere_interpolate(const c8mi_moliere_cubic* p, double x) {
  double inverse_step = (double)C8MI_MOLIERE_INTERVALS / (MOLIERE_TABLE_UPPER - 0.);
  double coordinate = (x - 0.) * inverse_step;
  if (coordinate < 0.) coordinate = 0.;
  if (coordinate > (double)C8MI_MOLIERE_INTERVALS) coordinate = (double)C8MI_MOLIERE_INTERVALS;
  uint32_t interval = (uint32_t)coordinate;
  if (interval >= C8MI_MOLIERE_INTERVALS) interval = C8MI_MOLIERE_INTERVALS - 1;
  double q = coordinate - (double)interval;
  const double* c = p[interval].c;
  return c[0] + q * (c[1] + q * (c[2] + q * c[3]));
}

/* f1_m / f2_m / integral_f1_m / integral_f2_m with a view: the table for x <= 20. */
static double mol_f1(const c8mi_moliere_snapshot* m, const c8mi_moliere_interpolation* t, double x) {
  return t && x <= MOLIERE_TABLE_UPPER ? moliere_interpolate(t->polynomials, x) : oracle_moliere_f1(m, x);
}
static double mol_f2(const c8mi_moliere_snapshot* m, const c8mi_moliere_interpolation* t, double x) {
  return t && x <= MOLIERE_TABLE_UPPER ? moliere_interpolate(t->polynomials + C8MI_MOLIERE_INTERVALS, x)
                                       : oracle_moliere_f2(m, x);
}
static double mol_if1(const c8mi_moliere_snapshot* m, const c8mi_moliere_interpolation* t, double x) {
  return t && x <= MOLIERE_TABLE_UPPER ? moliere_interpolate(t->polynomials + 2 * C8MI_MOLIERE_INTERVALS, x)
                                       : oracle_moliere_integral_f1(m, x);
}
static double mol_if2(const c8mi_moliere_snapshot* m, const c8mi_moliere_interpolation* t, double x) {
  return t && x <= MOLIERE_TABLE_UPPER ? moliere_interpolate(t->polynomials + 3 * C8MI_MOLIERE_INTERVALS, x)
                                       : oracle_moliere_integral_f2(m, x);
}

void oracle_moliere_interpolated_functions(const c8mi_moliere_snapshot* m, const c8mi_moliere_interpolation* t,
                                           double x, double out[4]) {
  out[0] = mol_f1(m, t, x);
  out[1] = mol_f2(m, t, x);
  out[2] = mol_if1(m, t, x);
  out[3] = mol_if2(m, t, x);
}

/* improve_moliere_initial_guess: the Gaussian coordinate plus the trilinear
 * interpolation of the tabulated correction in (reference B, beta^2, |coordinate|). */
static double moliere_improve_guess(const c8mi_moliere_interpolation* t, double reference_b, double beta2,
                                    double gaussian) {
  const double b_lo = 4.5, b_hi = 64., beta_lo = 0.25, beta_hi = 1., c_lo = 0., c_hi = 5.;
  const uint32_t nb = C8MI_MOLIERE_GUESS_B_NODES, nt = C8MI_MOLIERE_GUESS_BETA2_NODES,
                 nc = C8MI_MOLIERE_GUESS_COORDINATE_NODES;
  const double inv_b = (double)(nb - 1) / (b_hi - b_lo), inv_beta = (double)(nt - 1) / (beta_hi - beta_lo),
               inv_c = (double)(nc - 1) / (c_hi - c_lo);
  if (!t || !finite(reference_b) || !finite(beta2) || !finite(gaussian)) return gaussian;
  double mag = fabs(gaussian);
  if (reference_b < b_lo || reference_b > b_hi || beta2 < beta_lo || beta2 > beta_hi || mag < c_lo || mag > c_hi)
    return gaussian;
  double bc = (reference_b - b_lo) * inv_b, cc = (mag - c_lo) * inv_c, tc = (beta2 - beta_lo) * inv_beta;
  uint32_t bi = (uint32_t)bc, ti = (uint32_t)tc, ci = (uint32_t)cc;
  if (bi + 1 >= nb) bi = nb - 2;
  if (ti + 1 >= nt) ti = nt - 2;
  if (ci + 1 >= nc) ci = nc - 2;
  double bf = bc - (double)bi, tf = tc - (double)ti, cf = cc - (double)ci;
  const double* g = t->initial_guess_delta;
  size_t beta_stride = (size_t)nt * nc;
  size_t lb = (size_t)bi * beta_stride, ub = lb + beta_stride;
  size_t lt = (size_t)ti * nc + ci, ut = lt + nc;
  double ll = g[lb + lt] * (1. - cf) + g[lb + lt + 1] * cf;
  double lu = g[lb + ut] * (1. - cf) + g[lb + ut + 1] * cf;
  double ul = g[ub + lt] * (1. - cf) + g[ub + lt + 1] * cf;
  double uu = g[ub + ut] * (1. - cf) + g[ub + ut + 1] * cf;
  double lower = ll * (1. - tf) + lu * tf;
  double upper = ul * (1. - tf) + uu * tf;
  double corrected = mag + lower * (1. - bf) + upper * bf;
  if (!finite(corrected) || corrected < 0.) return gaussian;
  return gaussian < 0. ? -corrected : corrected;
}

static int moliere_build(const c8mi_moliere_snapshot* m, double grammage, double e_mev, mol_state_t* st) {
  if (m->component_count == 0 || m->component_count > C8MI_MOLIERE_MAX_COMPONENTS) return MS_INVALID_SNAPSHOT;
  if (!moliere_valid(m)) return MS_INVALID_SNAPSHOT;
  if (!finite(grammage) || grammage < 0. || !finite(e_mev) || !(e_mev > m->particle_mass_mev))
    return MS_INVALID_INPUT;
  if (grammage == 0.) return MS_NO_DEFLECTION;
  double mass = m->particle_mass_mev;
  double p2 = (e_mev - mass) * (e_mev + mass);
  double beta2 = 1. / (1. + mass * mass / p2);
  st->beta_squared = beta2;
  double bp2 = p2 / e_mev;
  bp2 *= bp2;
  double chi_c2 = 4. * MOLIERE_PI * m->avogadro_per_mol * m->fine_structure_constant *
                  m->fine_structure_constant * m->hbar_mev_s * m->hbar_mev_s * m->speed_of_light_cm_per_s *
                  m->speed_of_light_cm_per_s * grammage / bp2 * m->z_squared_over_a_average;
  if (!finite(chi_c2) || !(chi_c2 > 0.)) return MS_NON_FINITE;
  double b_max = 0.;
  for (uint32_t i = 0; i < m->component_count; ++i) {
    double chi_a2 = m->components[i].chi0_squared_mev2 / p2 * (1.13 + m->components[i].coulomb_correction / beta2);
    double b = oracle_moliere_solve_b(log(chi_c2 / chi_a2) + 1. - 2. * m->euler_mascheroni);
    if (!(b > 0.)) return MS_NO_DEFLECTION;
    double scale = chi_c2 * b;
    if (!finite(scale) || !(scale > 0.)) return MS_NON_FINITE;
    st->inverse_scale[i] = 1. / scale;
    st->inverse_b[i] = 1. / b;
    st->density_weight[i] = m->components[i].weight_zz / sqrt(scale * MOLIERE_PI);
    if (!finite(st->inverse_scale[i]) || !(st->inverse_scale[i] > 0.) || !finite(st->inverse_b[i]) ||
        !(st->inverse_b[i] > 0.) || !finite(st->density_weight[i]) || !(st->density_weight[i] > 0.))
      return MS_NON_FINITE;
    if (i == m->maximum_weight_index) b_max = b;
  }
  st->prefactor = sqrt(chi_c2 * b_max);
  return finite(st->prefactor) && st->prefactor > 0. ? MS_SUCCESS : MS_NON_FINITE;
}

static void moliere_evaluate(const c8mi_moliere_snapshot* m, const c8mi_moliere_interpolation* t,
                             const mol_state_t* st, double theta, double* density, double* cumulative) {
  double d = 0., cdf = 0., t2 = theta * theta;
  for (uint32_t i = 0; i < m->component_count; ++i) {
    double x = t2 * st->inverse_scale[i], ib = st->inverse_b[i], ib2 = ib * ib;
    d += st->density_weight[i] * (exp(-x) + mol_f1(m, t, x) * ib + mol_f2(m, t, x) * ib2);
    cdf += m->components[i].weight_zz *
           (0.5 * erf(sqrt(x > 0. ? x : 0.)) +
            sqrt(1. / MOLIERE_PI) * (mol_if1(m, t, x) * ib + mol_if2(m, t, x) * ib2));
  }
  *density = d * m->inverse_weight_zz_sum;
  cdf *= m->inverse_weight_zz_sum;
  *cumulative = theta < 0. ? -cdf : cdf;
}

static int moliere_one(const c8mi_moliere_snapshot* m, const c8mi_moliere_interpolation* t,
                       const mol_state_t* st, double u, uint32_t* iterations, double* angle) {
  *iterations = 0;
  *angle = 0.;
  if (!finite(u) || !(u > 0.) || !(u < 1.)) return MS_INVALID_INPUT;
  if (u == 0.5) return MS_SUCCESS;
  double gaussian = oracle_moliere_normal_ppf(u) / MOLIERE_SQRT2;
  double reference_b = 1. / st->inverse_b[m->maximum_weight_index];
  double current = st->prefactor * moliere_improve_guess(t, reference_b, st->beta_squared, gaussian);
  double target = u - 0.5;
  for (uint32_t it = 1; it <= 100; ++it) {
    double previous = current, density, cdf;
    moliere_evaluate(m, t, st, previous, &density, &cdf);
    *iterations = it;
    if (!finite(density) || density == 0.) {
      *angle = previous;
      return MS_NON_FINITE;
    }
    current = previous - (cdf - target) / density;
    *angle = current;
    if (!finite(current)) return MS_NON_FINITE;
    if (fabs((previous - current) / current) <= 5.e-5) return MS_SUCCESS;
  }
  *iterations = 100;
  return MS_NON_CONVERGENT;
}

/* Density and CDF of one projected angle for a step (test probe). */
int oracle_moliere_evaluate(const c8mi_moliere_snapshot* m, double grammage, double e_mev, double theta,
                            double* density, double* cumulative) {
  mol_state_t st;
  memset(&st, 0, sizeof st);
  int status = moliere_build(m, grammage, e_mev, &st);
  if (status != MS_SUCCESS) return status;
  moliere_evaluate(m, NULL, &st, theta, density, cumulative);
  return MS_SUCCESS;
}

/* Status, iteration count and the space angle of two projected draws; with a
 * table (may be NULL) as the reference's sampler with an interpolation view. */
int oracle_moliere_angle_interpolated(const c8mi_moliere_snapshot* m, const c8mi_moliere_interpolation* t,
                                      double grammage, double e_mev, double u1, double u2, uint32_t* iterations,
                                      double* angle) {
  mol_state_t st;
  memset(&st, 0, sizeof st);
  *iterations = 0;
  *angle = 0.;
  int status = moliere_build(m, grammage, e_mev, &st);
  if (status != MS_SUCCESS) return status;
  uint32_t i1, i2;
  double a1, a2;
  status = moliere_one(m, t, &st, u1, &i1, &a1);
  if (status != MS_SUCCESS) {
    *iterations = i1;
    *angle = a1;
    return status;
  }
  status = moliere_one(m, t, &st, u2, &i2, &a2);
  if (status != MS_SUCCESS) {
    *iterations = i2;
    *angle = a2;
    return status;
  }
  *iterations = i1 + i2;
  *angle = sqrt(a1 * a1 + a2 * a2);
  return finite(*angle) ? MS_SUCCESS : MS_NON_FINITE;
}

int oracle_moliere_angle(const c8mi_moliere_snapshot* m, double grammage, double e_mev, double u1, double u2,
                         uint32_t* iterations, double* angle) {
  return oracle_moliere_angle_interpolated(m, NULL, grammage, e_mev, u1, u2, iterations, angle);
}

static int moliere_rotate(const double d[3], double angle, double au, double out[3]) {
  if (!finite(angle) || angle < 0. || !finite(au) || !(au > 0.) || !(au < 1.)) return MS_INVALID_INPUT;
  double n2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const double lo = (1. - 1.e-12) * (1. - 1.e-12), hi = (1. + 1.e-12) * (1. + 1.e-12);
  if (!finite(n2) || n2 < lo || n2 > hi) return MS_INVALID_INPUT;
  double a1[3] = {0., 0., 0.};
  if (d[0] > 0.1) {
    a1[0] = -d[1];
    a1[1] = d[0];
  } else {
    a1[1] = -d[2];
    a1[2] = d[1];
  }
  double an = sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
  if (!finite(an) || !(an > 0.)) return MS_INVALID_INPUT;
  for (int k = 0; k < 3; ++k) a1[k] /= an;
  double ax[3] = {a1[1] * d[2] - a1[2] * d[1], a1[2] * d[0] - a1[0] * d[2], a1[0] * d[1] - a1[1] * d[0]};
  double ct = cos(angle), mst = -sin(angle);
  double z[3];
  for (int k = 0; k < 3; ++k) z[k] = d[k] * ct + ax[k] * mst;
  double az = au * 2. * MOLIERE_PI;
  double ca = cos(az), msa = -sin(az);
  double dz[3] = {d[1] * z[2] - d[2] * z[1], d[2] * z[0] - d[0] * z[2], d[0] * z[1] - d[1] * z[0]};
  double par = d[0] * z[0] + d[1] * z[1] + d[2] * z[2];
  for (int k = 0; k < 3; ++k) out[k] = z[k] * ca + dz[k] * msa + d[k] * par * (1. - ca);
  double on = sqrt(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
  if (!finite(on) || !(on > 0.)) return MS_NON_FINITE;
  for (int k = 0; k < 3; ++k) out[k] /= on;
  return MS_SUCCESS;
}

/* normalize(tracked + rotate(start) - start). */
int oracle_moliere_direction(const double start[3], const double tracked[3], double angle, double au,
                             double out[3]) {
  double scattered[3];
  int status = moliere_rotate(start, angle, au, scattered);
  if (status != MS_SUCCESS) {
    if (status == MS_NON_FINITE)
      for (int k = 0; k < 3; ++k) out[k] = scattered[k];
    return status;
  }
  double n2 = tracked[0] * tracked[0] + tracked[1] * tracked[1] + tracked[2] * tracked[2];
  const double lo = (1. - 1.e-12) * (1. - 1.e-12), hi = (1. + 1.e-12) * (1. + 1.e-12);
  if (!finite(n2) || n2 < lo || n2 > hi) return MS_INVALID_INPUT;
  double norm2 = 0.;
  for (int k = 0; k < 3; ++k) {
    out[k] = tracked[k] + scattered[k] - start[k];
    norm2 += out[k] * out[k];
  }
  double norm = sqrt(norm2);
  if (!finite(norm) || !(norm > 0.)) return MS_NON_FINITE;
  for (int k = 0; k < 3; ++k) out[k] /= norm;
  return MS_SUCCESS;
}

/* ========================== lepton step ========================== */

typedef struct {
  inter_t in;
  ps_t start, end;
  int limit;
  double continuous_deposit;
  double cut_deposit;
  int observed_before_cut;
  double traversed;   /* chord grammage of the step (g/cm^2) */
  int zero_length;    /* particle cut at the start: no scattering stage */
  int native_cut;     /* native_muon_pair_cut_record: signed deposit, no track */
} ltrans_t;

/* prepare_lepton_continuous_step + transport_lepton (LeptonContinuousStep.inl:43-120,
 * LeptonTransportStep.inl:27-513) in a field-free environment; the Molière stage
 * follows in lepton_source. */
static void transport_lepton(const ctx_t* x, const inter_t* in, ltrans_t* out, fb_t* fb) {
  const c8mi_environment* env = x->env;
  const cont_t* c = &x->cont;
  ps_t start = in->particle;
  memset(out, 0, sizeof *out);
  if (!is_charged_lepton(start.pid)) {
    set_fallback(fb, in, FR_UNSUPPORTED_PARTICLE);
    return;
  }
  if (in->status != IS_SELECTED && in->status != IS_DISTANCE_SAMPLED &&
      in->status != IS_NO_DISCRETE && in->status != IS_PARTICLE_CUT) {
    set_fallback(fb, in, FR_INVALID_FINAL_STATE);
    return;
  }
  if ((in->status == IS_SELECTED || in->status == IS_DISTANCE_SAMPLED) &&
      (!finite(in->interaction_grammage) || in->interaction_grammage < 0.)) {
    set_fallback(fb, in, FR_INVALID_FINAL_STATE);
    return;
  }
  if (in->status == IS_PARTICLE_CUT && native_muon_pair_cut(&start)) {
    /* native_muon_pair_cut_record (NativeMuonPairCutRecord.inl:12-25) */
    out->in = *in;
    out->in.vertex_reached = 0;
    out->start = out->end = start;
    out->end.step_id++;
    out->limit = LL_CUT;
    out->cut_deposit = start.energy_gev - transport_mass_gev(start.pid);
    out->zero_length = 1;
    out->native_cut = 1;
    return;
  }
  /* prepare_lepton_continuous_step */
  tq_t mass = cq_transport_mass(c, start.pid);
  tq_t minimum = cq_minimum_energy(c, start.pid);
  if (mass.status != TS_SUCCESS) {
    set_fallback(fb, in, fallback_reason(mass.status));
    return;
  }
  if (minimum.status != TS_SUCCESS) {
    set_fallback(fb, in, fallback_reason(minimum.status));
    return;
  }
  double e0 = start.energy_gev * 1000.;
  if (!(e0 > mass.value)) {
    set_fallback(fb, in, FR_INVALID_FINAL_STATE);
    return;
  }
  tq_t tcut = cq_transport_cut(c, start.pid);
  if (tcut.status != TS_SUCCESS || !finite(tcut.value) || !(tcut.value > 0.)) {
    set_fallback(fb, in, FR_INVALID_TABLE_QUERY);
    return;
  }
  /* muons with continuous tables stay with the host: no muon decay or muon
   * scattering in this restatement */
  if (!is_ee(start.pid)) {
    set_fallback(fb, in, FR_UNSUPPORTED_PARTICLE);
    return;
  }
  int particle_cut = in->status == IS_PARTICLE_CUT || e0 - mass.value < tcut.value;
  double max_grammage = 0.;
  if (!particle_cut) {
    tq_t r0 = cq_range(c, start.pid, e0);
    if (r0.status != TS_SUCCESS) {
      set_fallback(fb, in, fallback_reason(r0.status));
      return;
    }
    double rl = (1. - MAX_CONTINUOUS_RELATIVE_LOSS) * e0;
    double target = rl > minimum.value ? rl : minimum.value;
    tq_t r1 = cq_range(c, start.pid, target);
    if (r1.status != TS_SUCCESS) {
      set_fallback(fb, in, fallback_reason(r1.status));
      return;
    }
    max_grammage = r0.value - r1.value;
    double scale = r0.value > 1. ? r0.value : 1.;
    if (max_grammage < 0. && max_grammage > -64. * 2.22044604925031308085e-16 * scale)
      max_grammage = 0.;
    if (!finite(max_grammage) || max_grammage < 0.) {
      set_fallback(fb, in, FR_INVALID_TABLE_QUERY);
      return;
    }
  }
  tq_t native_mass = cq_mass(c, start.pid);
  if (particle_cut) {
    layer_t layer = query_environment(env, start.position_m, start.direction, start.environment_node);
    if (layer.status != ES_SUCCESS) {
      set_fallback(fb, in, FR_UNSUPPORTED_GEOMETRY);
      return;
    }
    start.environment_node = layer.node;
    out->in = *in;
    out->in.vertex_reached = 0;
    out->in.particle_mass_gev = native_mass.value / 1000.;
    out->start = start;
    out->end = start;
    out->zero_length = 1;
    out->end.step_id++;
    out->limit = LL_CUT;
    out->cut_deposit = (e0 - mass.value) / 1000.;
    return;
  }
  layer_t layer = query_environment(env, start.position_m, start.direction, start.environment_node);
  if (layer.status != ES_SUCCESS) {
    set_fallback(fb, in, FR_UNSUPPORTED_GEOMETRY);
    return;
  }
  start.environment_node = layer.node;
  const c8mi_env_model* model = get_model(env, layer.node);
  double m_gev = mass.value / 1000.;
  /* maximum_uniform_magnetic_step, lepton_boundary, lepton_observation_distance
   * (LeptonTransportStep.inl:119-160) */
  double charge = start.pid > 0 ? -1. : 1.;
  const double* field = model->magnetic_field_t;
  mlimit_t mlimit = magnetic_step_limit(&start, m_gev, charge, field, env->maximum_magnetic_deflection_rad);
  if (mlimit.status == MAG_INVALID || mlimit.status == MAG_NONFINITE) {
    set_fallback(fb, in, FR_MAGNETIC_TRANSPORT);
    return;
  }
  double geometry;
  int32_t next_node;
  if (!magnetic_lepton_boundary(env, &start, m_gev, charge, field, &mlimit, &geometry, &next_node)) {
    set_fallback(fb, in, FR_MAGNETIC_BOUNDARY);
    return;
  }
  int magnetic_wins = next_node == start.environment_node;
  int obs_wins = 0;
  if (env->observation_enabled) {
    double o = magnetic_observation_distance(env, &start, m_gev, charge, field, &mlimit);
    if (o < geometry) {
      geometry = o;
      obs_wins = 1;
    }
    if (obs_wins) magnetic_wins = 0;
  }
  gq_t gg = competition_grammage(env, layer.node, start.position_m, start.direction, geometry);
  if (gg.status != ES_SUCCESS) {
    set_fallback(fb, in, FR_GRAMMAGE);
    return;
  }
  int has_interaction = in->status == IS_SELECTED || in->status == IS_DISTANCE_SAMPLED;
  double ig = has_interaction ? in->interaction_grammage : HUGE_VAL;
  int continuous_wins = max_grammage < gg.value && max_grammage < ig;
  int interaction_wins = !continuous_wins && has_interaction && ig <= gg.value;
  out->in = *in;
  out->in.vertex_reached = 0;
  out->in.particle_mass_gev = native_mass.value / 1000.;
  out->start = start;
  double traversed = gg.value;
  double distance = geometry;
  if (continuous_wins) traversed = max_grammage;
  else if (interaction_wins) traversed = ig;
  if (continuous_wins || interaction_wins) {
    gq_t d = environment_arclength(env, layer.node, start.position_m, start.direction, traversed);
    if (d.status != ES_SUCCESS || d.value > geometry * (1. + 1.e-12)) {
      set_fallback(fb, in, FR_INVERSE_GRAMMAGE);
      return;
    }
    distance = d.value;
  }
  int observation_reached = obs_wins && !continuous_wins && !interaction_wins && !magnetic_wins;
  /* advance_uniform_magnetic_field (UniformMagneticField.inl:146-195), no field
   * where the limit was Linear */
  const double zero_field[3] = {0., 0., 0.};
  int advanced = magnetic_advance(&start, m_gev, charge, mlimit.status == MAG_LINEAR ? zero_field : field, distance,
                                  &out->end);
  if (advanced == MAG_INVALID || advanced == MAG_NONFINITE) {
    set_fallback(fb, in, FR_MAGNETIC_TRANSPORT);
    return;
  }
  /* chord grammage (LeptonTransportStep.inl:301-331) */
  double chord_dir[3], chord2 = 0.;
  for (int a = 0; a < 3; ++a) {
    double dd = out->end.position_m[a] - start.position_m[a];
    chord_dir[a] = dd;
    chord2 += dd * dd;
  }
  double chord = sqrt(chord2);
  if (chord > 0.) {
    for (int a = 0; a < 3; ++a) chord_dir[a] /= chord;
  } else {
    for (int a = 0; a < 3; ++a) chord_dir[a] = start.direction[a];
  }
  gq_t cg = environment_grammage(env, layer.node, start.position_m, chord_dir, chord);
  if (cg.status != ES_SUCCESS) {
    set_fallback(fb, in, FR_GRAMMAGE);
    return;
  }
  traversed = cg.value;
  out->traversed = traversed;
  /* evaluate_lepton_continuous_loss (LeptonContinuousStep.inl:122-159) */
  tq_t fe = cq_energy_after_loss(c, start.pid, e0, traversed);
  double final_mev;
  int cut_reached = 0;
  if (fe.status == TS_SUCCESS) {
    final_mev = fe.value;
  } else if (fe.status == TS_TRANSPORT_CUT) {
    final_mev = minimum.value;
    cut_reached = 1;
  } else {
    set_fallback(fb, in, fallback_reason(fe.status));
    return;
  }
  double loss = e0 - final_mev;
  if (continuous_wins && !(loss > 0.)) {
    set_fallback(fb, in, FR_INVALID_TABLE_QUERY);
    return;
  }
  out->end.energy_gev = final_mev / 1000.;
  out->continuous_deposit = loss / 1000.;
  if (out->end.time_s > PARTICLE_CUT_TIME_S) {
    out->end.step_id++;
    out->limit = LL_CUT;
    out->cut_deposit = (final_mev - mass.value) / 1000.;
    out->observed_before_cut = observation_reached;
    return;
  }
  int reaches_cut = cut_reached || final_mev - mass.value < tcut.value;
  if (reaches_cut) {
    out->end.step_id++;
    out->limit = LL_CUT;
    if (!observation_reached) {
      layer_t el = query_environment(env, out->end.position_m, out->end.direction, start.environment_node);
      if (el.status != ES_SUCCESS) {
        set_fallback(fb, in, FR_UNSUPPORTED_GEOMETRY);
        return;
      }
    }
    out->cut_deposit = (final_mev - mass.value) / 1000.;
    out->observed_before_cut = observation_reached;
    return;
  }
  if (interaction_wins) {
    layer_t v = query_environment(env, out->end.position_m, out->end.direction, start.environment_node);
    if (v.status != ES_SUCCESS || !finite(v.density) || !(v.density > 0.)) {
      set_fallback(fb, in, FR_VERTEX_LOOKUP);
      return;
    }
    out->limit = LL_INTERACTION;
    out->in.particle = out->end;
    out->in.status = IS_RESELECT;
    out->in.vertex_reached = 1;
    out->in.process_id = 0;
    out->in.component_hash = 0;
    out->in.energy_fraction = 0.;
    out->in.loss_quantile = 0.;
    out->in.mass_density = v.density;
    return;
  }
  out->end.step_id++;
  if (continuous_wins) {
    layer_t el = query_environment(env, out->end.position_m, out->end.direction, start.environment_node);
    if (el.status != ES_SUCCESS) {
      set_fallback(fb, in, FR_UNSUPPORTED_GEOMETRY);
      return;
    }
    out->limit = LL_CONTINUOUS;
    return;
  }
  if (magnetic_wins) {
    layer_t el = query_environment(env, out->end.position_m, out->end.direction, start.environment_node);
    if (el.status != ES_SUCCESS) {
      set_fallback(fb, in, FR_MAGNETIC_BOUNDARY);
      return;
    }
    out->limit = LL_MAGNETIC;
    return;
  }
  if (observation_reached) {
    out->limit = LL_OBSERVATION;
    return;
  }
  out->end.environment_node = next_node;
  const c8mi_env_model* nm = get_model(env, next_node);
  if (!nm) {
    out->limit = LL_ESCAPED;
    return;
  }
  layer_t next = query_environment(env, out->end.position_m, out->end.direction, next_node);
  if (next.status != ES_SUCCESS || next_node == start.environment_node) {
    set_fallback(fb, in, FR_UNSUPPORTED_GEOMETRY);
    return;
  }
  out->limit = LL_LAYER;
  out->end.medium_id = nm->medium_id;
}

/* sample_annihilation_rho (LeptonFinalStateStep.inl:150-192) */
static double ann_term(double a1, double a2, double rho) { return a1 * log(rho) + a2 / rho - rho; }
static double ann_eq(double v, double a1, double a2, double lt, double ut, double u) {
  return ann_term(a1, a2, v) - lt - u * (ut - lt);
}
static int annihilation_rho(double e, double m, double u, double* rho) {
  double gamma = e / m;
  if (!finite(gamma) || !(gamma > 1.) || !finite(u) || !(u > 0.) || !(u < 1.)) return 0;
  double aux = sqrt((gamma - 1.) / (gamma + 1.));
  double lower = 0.5 * (1. - aux), upper = 0.5 * (1. + aux);
  double a2 = 1. / ((gamma + 1.) * (gamma + 1.));
  double a1 = 1. + 2. * gamma * a2;
  double lt = ann_term(a1, a2, lower), ut = ann_term(a1, a2, upper);
  double lv = ann_eq(lower, a1, a2, lt, ut, u), uv = ann_eq(upper, a1, a2, lt, ut, u);
  if (!finite(lv) || !finite(uv) || lv * uv > 0.) return 0;
  double precision = lower * (e * 1000.) * PROPOSAL_HALF_PRECISION;
  if (!finite(precision) || !(precision > 0.)) return 0;
  for (int it = 0; it <= 100; ++it) {
    double center = (lower + upper) / 2.;
    double cv = ann_eq(center, a1, a2, lt, ut, u);
    if (sgn(cv) == sgn(lv)) {
      lower = center;
      lv = cv;
    } else {
      upper = center;
    }
    if (fabs(upper - lower) < precision) {
      *rho = lower;
      return finite(*rho) && *rho > 0. && *rho < 1.;
    }
  }
  *rho = lower;
  return finite(*rho) && *rho > 0. && *rho < 1.;
}

typedef struct {
  ps_t next[2];
  uint32_t next_count;
  ps_t photons[2];
  uint32_t photon_count;
  uint32_t child_count;
} lout_t;

static void lepton_source(shower_t* s, const ps_t* p, lout_t* o, uint64_t child_offset,
                          uint64_t first_history) {
  const ctx_t* x = &s->x;
  const c8mi_em_config* cfg = x->cfg;
  memset(o, 0, sizeof *o);
  inter_t in;
  fb_t fb = {0, 0, {0}, 0};
  select_interaction(x, p, &in, &fb);
  if (fb.flag) {
    record_fallback(s, &fb, STAGE_SELECTION);
    return;
  }
  ltrans_t tr;
  transport_lepton(x, &in, &tr, &fb);
  if (fb.flag) {
    record_fallback(s, &fb, STAGE_TRANSPORT);
    return;
  }
  /* apply_moliere_scattering_stage (MoliereStep.inl:74-99): draws keyed by the
   * start state; the end direction becomes normalize(tracked + scattered - start) */
  if (cfg->moliere_enabled && !tr.zero_length) {
    const ps_t* st = &tr.start;
    double u1 = uniform(x->seed, x->shower, st, RP_SCATTERING, 0);
    double u2 = uniform(x->seed, x->shower, st, RP_SCATTERING, 1);
    double ua = uniform(x->seed, x->shower, st, RP_SCATTERING, 2);
    uint32_t its;
    double angle;
    int ms = oracle_moliere_angle_interpolated(&cfg->electron_moliere, cfg->electron_moliere_interpolation,
                                               tr.traversed, st->energy_gev * 1000., u1, u2, &its, &angle);
    if (ms != MS_NO_DEFLECTION) {
      double dir[3];
      if (ms != MS_SUCCESS ||
          oracle_moliere_direction(tr.start.direction, tr.end.direction, angle, ua, dir) != MS_SUCCESS) {
        set_fallback(&fb, &tr.in, FR_MOLIERE_FAILED);
        record_fallback(s, &fb, STAGE_TRANSPORT);
        return;
      }
      for (int k = 0; k < 3; ++k) tr.end.direction[k] = dir[k];
    }
    if (tr.limit == LL_INTERACTION) tr.in.particle = tr.end;
  }
  s->stats.lepton_steps++;
  double w = tr.start.weight;
  if (tr.native_cut) {
    /* the native cut's profile step (ProfileAccumulation.inl:300-345): no
     * track, the signed deposit to the muon energy loss, the rest mass */
    double deposit = tr.cut_deposit * w;
    if (deposit < 0. && cfg->profile_bins != 0) s->prof.counters[OR_EMC_DEPOSITED_STEPS]++;
    ledger_add(&s->prof, cfg, OR_EMC_NATIVE_MUON_CUT, deposit);
    s->prof.counters[OR_EMC_STEPS]++;
    s->prof.counters[OR_EMC_LEPTON_LIMIT0 + LL_CUT]++;
    ledger_add(&s->prof, cfg, OR_EMC_CUT_REST_MASS, transport_mass_gev(tr.start.pid) * w);
    return;
  }
  /* the front's step record feeds the radio accumulator
   * (LeptonCascade.inl:81-82) */
  record_track(s, &tr.start, &tr.end);
  /* accumulate_lepton_profile_step (ProfileAccumulation.inl:300-375) */
  double gs = project_grammage(cfg, tr.start.position_m);
  double ge = project_grammage(cfg, tr.end.position_m);
  particle_profile(&s->prof, cfg, tr.start.pid, gs, ge, w);
  energy_profile(&s->prof, cfg, gs, ge, tr.continuous_deposit * w);
  point_energy_profile(&s->prof, cfg, ge, tr.cut_deposit * w);
  ledger_add(&s->prof, cfg, OR_EMC_CONTINUOUS, tr.continuous_deposit * w);
  s->prof.counters[OR_EMC_STEPS]++;
  s->prof.counters[OR_EMC_LEPTON_LIMIT0 + tr.limit]++;
  double tm = transport_mass_gev(tr.start.pid);
  if (tr.limit == LL_CUT) {
    ledger_add(&s->prof, cfg, OR_EMC_CUT_DEPOSIT, tr.cut_deposit * w);
    ledger_add(&s->prof, cfg, OR_EMC_CUT_REST_MASS, tm * w);
    if (tr.observed_before_cut) {
      ledger_add(&s->prof, cfg, OR_EMC_OVERLAP, (tr.cut_deposit + tm) * w);
      terminal_energy(s, tr.end.energy_gev * w, 1);
      record_observation(s, &tr.end);
    }
    return;
  }
  if (tr.limit == LL_OBSERVATION || tr.limit == LL_ESCAPED) {
    terminal_energy(s, tr.end.energy_gev * w, tr.limit == LL_OBSERVATION);
    record_observation(s, &tr.end);
    return;
  }
  if (tr.limit != LL_INTERACTION) {
    o->next[o->next_count++] = tr.end;
    return;
  }
  /* select_lepton_vertex (LeptonVertexSelection.inl:41-197) */
  s->stats.lepton_interactions++;
  inter_t r = tr.in;
  const ps_t* vp = &r.particle;
  fb_t f;
  if (r.status != IS_RESELECT || !finite(r.total_rate) || !(r.total_rate > 0.) ||
      !finite(r.process_uniform) || !(r.process_uniform > 0.) || !(r.process_uniform < 1.)) {
    set_fallback(&f, &r, FR_INVALID_FINAL_STATE);
    record_fallback(s, &f, STAGE_VERTEX);
    return;
  }
  double e_mev = vp->energy_gev * 1000.;
  double outer = r.process_uniform * r.total_rate;
  tq_t vt = dq_total_rate(&x->cont, vp->pid, e_mev);
  if (vt.status != TS_SUCCESS) {
    set_fallback(&f, &r, fallback_reason(vt.status));
    record_fallback(s, &f, STAGE_VERTEX);
    return;
  }
  {
    or_vertex* lv = log_vertex(s, 0, vp);
    if (lv) {
      lv->accepted = vt.value > 0. && outer < vt.value;
      lv->rate_start = r.total_rate;
      lv->rate_vertex = vt.value;
      lv->process_uniform = r.process_uniform;
    }
  }
  if (!(vt.value > 0.) || outer >= vt.value) {
    if (vp->step_id == 0xffffffffffffffffULL) {
      set_fallback(&f, &r, FR_INVALID_FINAL_STATE);
      record_fallback(s, &f, STAGE_VERTEX);
      return;
    }
    ps_t cont = *vp;
    cont.step_id++;
    o->next[o->next_count++] = cont;
    return;
  }
  double us = uniform(x->seed, x->shower, vp, RP_SELECTION, 0);
  sel_t sel = select_column(&x->cont, vp->pid, e_mev, us);
  if (sel.status != TS_SUCCESS) {
    r.process_id = sel.process;
    set_fallback(&f, &r, fallback_reason(sel.status));
    record_fallback(s, &f, STAGE_VERTEX);
    return;
  }
  if (!sel.selected) {
    set_fallback(&f, &r, FR_ZERO_TOTAL_RATE);
    record_fallback(s, &f, STAGE_VERTEX);
    return;
  }
  r.process_id = sel.process;
  r.component_hash = sel.component;
  r.loss_quantile = sel.quantile;
  if (requires_replay(vp->pid, sel.process, sel.quantile)) {
    set_fallback(&f, &r, FR_SELECTION_REPLAY);
    record_fallback(s, &f, STAGE_VERTEX);
    return;
  }
  tq_t loss = dq_loss_fraction(&x->cont, vp->pid, sel.process, sel.component, e_mev, sel.quantile);
  if (loss.status != TS_SUCCESS) {
    set_fallback(&f, &r, fallback_reason(loss.status));
    record_fallback(s, &f, STAGE_VERTEX);
    return;
  }
  r.status = IS_SELECTED;
  r.energy_fraction = loss.value;

  /* classify_lepton_final_state (LeptonFinalStateStep.inl:322-498), e+- only */
  const ps_t* parent = &r.particle;
  int cap = capability(parent->pid, r.process_id);
  if (cap != CAP_BREMS && cap != CAP_ANNIHILATION && cap != CAP_IONIZATION) {
    set_fallback(&f, &r, cap == CAP_CPU_ONLY ? FR_CPU_ONLY : FR_NOT_IMPLEMENTED);
    record_fallback(s, &f, STAGE_FINAL);
    return;
  }
  const c8mi_brems_lpm* lpm = &cfg->brems_lpm;
  double v = r.energy_fraction;
  double me = lpm->electron_mass_mev / 1000.;
  double lm = r.particle_mass_gev > 0. ? r.particle_mass_gev : lpm->lepton_mass_mev / 1000.;
  double fs_me = lpm->lepton_mass_mev / 1000.; /* materialisation electron mass */
  if (!finite(parent->energy_gev) || !(parent->energy_gev > lm) || parent->generation == 0xffffffffU ||
      parent->step_id == 0xffffffffffffffffULL) {
    set_fallback(&f, &r, FR_INVALID_FINAL_STATE);
    record_fallback(s, &f, STAGE_FINAL);
    return;
  }
  if (cap == CAP_ANNIHILATION) {
    if (!finite(v) || fabs(v - 1.) > 1.e-12) {
      set_fallback(&f, &r, FR_INVALID_FINAL_STATE);
      record_fallback(s, &f, STAGE_FINAL);
      return;
    }
    double ru = uniform(x->seed, x->shower, parent, P_ANNIHILATION, D_ANNIHILATION_RHO);
    double au = uniform(x->seed, x->shower, parent, P_ANNIHILATION, D_ANNIHILATION_AZIMUTH);
    double rho;
    if (!annihilation_rho(parent->energy_gev, lm, ru, &rho)) {
      set_fallback(&f, &r, FR_INVALID_FINAL_STATE);
      record_fallback(s, &f, STAGE_FINAL);
      return;
    }
    double total = parent->energy_gev + lm;
    thin_t th;
    if (!two_child_thinning(x, parent, P_ANNIHILATION, total * (1. - rho), total * rho, 1, &th)) {
      set_fallback(&f, &r, FR_INVALID_FINAL_STATE);
      record_fallback(s, &f, STAGE_FINAL);
      return;
    }
    /* materialize (LeptonFinalStateStep.inl:700-790), in flight */
    double e1 = total * (1. - rho), e2 = total * rho;
    double mom = sqrt((parent->energy_gev + lm) * (parent->energy_gev - lm));
    double c1 = (total * (1. - rho) - lm) / ((1. - rho) * mom);
    double c2 = (total * rho - lm) / (rho * mom);
    if (!finite(e1) || !finite(e2) || !(e1 > 0.) || !(e2 > 0.) || !finite(c1) || !finite(c2)) {
      s->prof.counters[OR_EMC_INVALID]++;
      return;
    }
    double az = au * LEPTON_TWO_PI;
    ps_t g1 = *parent;
    g1.pid = 22;
    g1.energy_gev = e1;
    g1.parent_history_id = parent->history_id;
    g1.generation = parent->generation + 1;
    g1.step_id = 0;
    g1.reserved = 0;
    g1.weight = th.w1;
    ps_t g2 = g1;
    g2.energy_gev = e2;
    g2.weight = th.w2;
    deflect(parent->direction, c1, az, g1.direction, 1);
    deflect(parent->direction, c2, fmod(az + LEPTON_PI, LEPTON_TWO_PI), g2.direction, 1);
    {
      or_vertex* lv = log_vertex(s, 1, parent);
      if (lv) {
        lv->process = r.process_id;
        lv->mass = lm;
        lv->target_mass = lm;
        lv->v = v;
        log_child(lv, 22, e1, g1.direction);
        log_child(lv, 22, e2, g2.direction);
      }
      log_thinning(lv, (uint32_t)th.status, th.keep, parent->weight, th.w1, th.w2);
    }
    capture_first(s, parent, P_ANNIHILATION, &g1, &g2);
    uint32_t k = 0;
    if (th.keep & 1U) {
      g1.history_id = first_history + child_offset + k++;
      o->photons[o->photon_count++] = g1;
    }
    if (th.keep & 2U) {
      g2.history_id = first_history + child_offset + k++;
      o->photons[o->photon_count++] = g2;
    }
    o->child_count = k;
    s->stats.secondaries += k;
    ledger_add(&s->prof, cfg, OR_EMC_MEDIUM_REST_MASS, ELECTRON_MASS_GEV * w);
    thinning_ledger(s, th.status, 2, k, (e1 + e2) * w,
                    ((th.keep & 1U) ? e1 * th.w1 : 0.) + ((th.keep & 2U) ? e2 * th.w2 : 0.));
    return;
  }
  if (cap == CAP_IONIZATION) {
    double ev = (v * (parent->energy_gev * 1000.)) / (parent->energy_gev * 1000.);
    double outgoing = parent->energy_gev * (1. - ev);
    double delta = parent->energy_gev * ev + me;
    if (!finite(ev) || !(ev > 0.) || !(ev < 1.) || !finite(outgoing) || outgoing < lm ||
        !finite(delta) || !(delta > me)) {
      set_fallback(&f, &r, FR_INVALID_FINAL_STATE);
      record_fallback(s, &f, STAGE_FINAL);
      return;
    }
    double au = uniform(x->seed, x->shower, parent, P_IONIZATION, D_IONIZATION_AZIMUTH);
    thin_t th;
    double e1 = secondary_transport_energy(parent->pid, outgoing, lm);
    double e2 = secondary_transport_energy(11, delta, me);
    if (!two_child_thinning(x, parent, P_IONIZATION, e1, e2, 1, &th)) {
      set_fallback(&f, &r, FR_INVALID_FINAL_STATE);
      record_fallback(s, &f, STAGE_FINAL);
      return;
    }
    /* materialize (LeptonFinalStateStep.inl:796-876) */
    double mv = (r.energy_fraction * (parent->energy_gev * 1000.)) / (parent->energy_gev * 1000.);
    double oe = parent->energy_gev * (1. - mv);
    double de = parent->energy_gev * mv + fs_me;
    double im = sqrt((parent->energy_gev + lm) * (parent->energy_gev - lm));
    double om = sqrt((oe + lm) * (oe - lm));
    double dm = sqrt((de + fs_me) * (de - fs_me));
    double oc = ((parent->energy_gev + fs_me) * oe - parent->energy_gev * fs_me - lm * lm) / (im * om);
    double dc = ((parent->energy_gev + fs_me) * de - parent->energy_gev * fs_me - fs_me * fs_me) / (im * dm);
    if (!finite(oc) || !finite(dc)) {
      s->prof.counters[OR_EMC_INVALID]++;
      return;
    }
    double az = au * LEPTON_TWO_PI;
    ps_t out = *parent;
    out.energy_gev = secondary_transport_energy(out.pid, oe, lm);
    out.parent_history_id = parent->history_id;
    out.generation = parent->generation + 1;
    out.step_id = 0;
    out.reserved = 0;
    out.weight = th.w1;
    ps_t dl = out;
    dl.pid = 11;
    dl.energy_gev = secondary_transport_energy(11, de, fs_me);
    dl.weight = th.w2;
    deflect(parent->direction, oc, az, out.direction, 1);
    deflect(parent->direction, dc, fmod(az + LEPTON_PI, LEPTON_TWO_PI), dl.direction, 1);
    {
      or_vertex* lv = log_vertex(s, 1, parent);
      if (lv) {
        lv->process = r.process_id;
        lv->mass = lm;
        lv->target_mass = fs_me;
        lv->v = r.energy_fraction;
        log_child(lv, out.pid, oe, out.direction);
        log_child(lv, 11, de, dl.direction);
      }
      log_thinning(lv, (uint32_t)th.status, th.keep, parent->weight, th.w1, th.w2);
    }
    capture_first(s, parent, P_IONIZATION, &out, &dl);
    uint32_t k = 0;
    if (th.keep & 1U) {
      out.history_id = first_history + child_offset + k++;
      o->next[o->next_count++] = out;
    }
    if (th.keep & 2U) {
      dl.history_id = first_history + child_offset + k++;
      o->next[o->next_count++] = dl;
    }
    o->child_count = k;
    s->stats.secondaries += k;
    ledger_add(&s->prof, cfg, OR_EMC_MASS_CORRECTION,
               mass_correction(parent->pid, lm, (th.keep & 1U) ? th.w1 : 0.) +
                   mass_correction(11, fs_me, (th.keep & 2U) ? th.w2 : 0.));
    ledger_add(&s->prof, cfg, OR_EMC_MEDIUM_REST_MASS, ELECTRON_MASS_GEV * w);
    thinning_ledger(s, th.status, 2, k, (out.energy_gev + dl.energy_gev) * w,
                    ((th.keep & 1U) ? out.energy_gev * th.w1 : 0.) +
                        ((th.keep & 2U) ? dl.energy_gev * th.w2 : 0.));
    return;
  }
  /* bremsstrahlung */
  double pe = parent->energy_gev * v;
  double le = parent->energy_gev - pe;
  if (!finite(v) || !(v > 0.) || !(v < 1.) || !finite(pe) || !(pe > 0.) || !finite(le) || le < lm) {
    set_fallback(&f, &r, FR_INVALID_FINAL_STATE);
    record_fallback(s, &f, STAGE_FINAL);
    return;
  }
  double au = uniform(x->seed, x->shower, parent, P_BREMS, D_BREMS_AZIMUTH);
  double survival = 0.;
  int ls = brems_lpm(lpm, r.component_hash, parent->energy_gev * 1000., v, r.mass_density, &survival);
  if (ls != 0) {
    set_fallback(&f, &r, ls == 3 ? FR_INVALID_DENSITY : FR_LPM_UNAVAILABLE);
    record_fallback(s, &f, STAGE_FINAL);
    return;
  }
  double lu = uniform(x->seed, x->shower, parent, P_BREMS, D_BREMS_LPM);
  if (lu > survival) {
    ps_t cont = *parent;
    cont.step_id++;
    o->next[o->next_count++] = cont;
    return;
  }
  thin_t th;
  double e1 = secondary_transport_energy(parent->pid, le, lm);
  if (!two_child_thinning(x, parent, P_BREMS, e1, pe, 1, &th)) {
    set_fallback(&f, &r, FR_INVALID_FINAL_STATE);
    record_fallback(s, &f, STAGE_FINAL);
    return;
  }
  /* materialize (LeptonFinalStateStep.inl:957-1025) */
  double az = au * LEPTON_TWO_PI;
  double pcos = cos(lm / parent->energy_gev);
  double pdir[3];
  deflect(parent->direction, pcos, az, pdir, 1);
  ps_t lep = *parent;
  lep.energy_gev = e1;
  lep.parent_history_id = parent->history_id;
  lep.generation = parent->generation + 1;
  lep.step_id = 0;
  lep.reserved = 0;
  lep.weight = th.w1;
  double pm2 = (parent->energy_gev + lm) * (parent->energy_gev - lm);
  double pmom = sqrt(0. > pm2 ? 0. : pm2);
  for (int a = 0; a < 3; ++a) lep.direction[a] = parent->direction[a] * pmom - pdir[a] * pe;
  double n2 = dot3(lep.direction, lep.direction);
  if (!finite(n2) || !(n2 > 0.)) {
    s->prof.counters[OR_EMC_INVALID]++;
    return;
  }
  double inv = 1. / sqrt(n2);
  for (int a = 0; a < 3; ++a) lep.direction[a] *= inv;
  if (!finite(lep.direction[0]) || !finite(lep.direction[1]) || !finite(lep.direction[2])) {
    s->prof.counters[OR_EMC_INVALID]++;
    return;
  }
  ps_t ph = *parent;
  ph.pid = 22;
  ph.energy_gev = pe;
  ph.parent_history_id = parent->history_id;
  ph.generation = parent->generation + 1;
  ph.step_id = 0;
  ph.reserved = 0;
  ph.weight = th.w2;
  for (int a = 0; a < 3; ++a) ph.direction[a] = pdir[a];
  {
    or_vertex* lv = log_vertex(s, 1, parent);
    if (lv) {
      lv->process = r.process_id;
      lv->mass = lm;
      lv->v = v;
      log_child(lv, parent->pid, le, lep.direction);
      log_child(lv, 22, pe, pdir);
    }
    log_thinning(lv, (uint32_t)th.status, th.keep, parent->weight, th.w1, th.w2);
  }
  capture_first(s, parent, P_BREMS, &lep, &ph);
  uint32_t k = 0;
  if (th.keep & 1U) {
    lep.history_id = first_history + child_offset + k++;
    o->next[o->next_count++] = lep;
  }
  if (th.keep & 2U) {
    ph.history_id = first_history + child_offset + k++;
    o->photons[o->photon_count++] = ph;
  }
  o->child_count = k;
  s->stats.secondaries += k;
  ledger_add(&s->prof, cfg, OR_EMC_MASS_CORRECTION,
             mass_correction(parent->pid, lm, (th.keep & 1U) ? th.w1 : 0.));
  thinning_ledger(s, th.status, 2, k, (e1 + pe) * w,
                  ((th.keep & 1U) ? e1 * th.w1 : 0.) + ((th.keep & 2U) ? pe * th.w2 : 0.));
}

/* ========================== driver ========================== */

/* wavefront_bucket_key (WavefrontBucketing.inl:24-79): PID bucket (photon 0,
 * e- 1, e+ 2, other 3) << 48 | (medium ^ 2^31) << 16 | energy bin, the bin
 * floor(log2(E/GeV) * 16 + 32768) clamped to [0, 65534], 65535 for a
 * non-positive or non-finite energy. */
uint64_t oracle_em_bucket_key(const c8mi_particle_state* p) {
  uint64_t const pid = p->pid == 22 ? 0u : p->pid == 11 ? 1u : p->pid == -11 ? 2u : 3u;
  uint64_t const medium = (uint64_t)((uint32_t)p->medium_id ^ 0x80000000u);
  uint64_t bin;
  double const e = p->energy_gev;
  if (!(e > 0.) || !finite(e)) {
    bin = 65535u;
  } else {
    double const c = floor(log2(e) * 16. + 32768.);
    bin = !(c > 0.) ? 0u : c >= 65534. ? 65534u : (uint64_t)c;
  }
  return pid << 48 | medium << 16 | bin;
}

typedef struct {
  uint64_t key;
  uint64_t index;
} sort_item_t;

static int sort_item_less(const void* a, const void* b) {
  const sort_item_t* x = (const sort_item_t*)a;
  const sort_item_t* y = (const sort_item_t*)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->index < y->index ? -1 : x->index > y->index;
}

/* Stable reorder of a queue by key(i): ties keep the queue order. */
static void stable_reorder(queue_t* q, uint64_t (*key)(const queue_t*, uint64_t)) {
  uint64_t const n = q->count;
  sort_item_t* order = (sort_item_t*)malloc(n * sizeof *order);
  ps_t* items = (ps_t*)malloc(n * sizeof *items);
  uint8_t* cls = (uint8_t*)malloc(n);
  for (uint64_t i = 0; i < n; ++i) {
    order[i].key = key(q, i);
    order[i].index = i;
  }
  qsort(order, n, sizeof *order, sort_item_less);
  for (uint64_t j = 0; j < n; ++j) {
    items[j] = q->items[order[j].index];
    cls[j] = q->cls[order[j].index];
  }
  free(q->items);
  free(q->cls);
  q->items = items;
  q->cls = cls;
  q->capacity = n;
  free(order);
}

static uint64_t class_key(const queue_t* q, uint64_t i) { return q->cls[i]; }
static uint64_t bucket_key(const queue_t* q, uint64_t i) { return oracle_em_bucket_key(&q->items[i]); }

/* One front of one species over the whole queue, in queue order (after the
 * bucketing of PhotonCascade.inl:318-332 / LeptonCascade.inl:296-307 for
 * fronts of at least 256 particles): continuing particles of the species form
 * the next front, the other species' particles join its queue as pending;
 * children take history IDs from the call's running counter in source order. */
static void bucket_front(shower_t* s, queue_t* q) {
  if (q->count >= BUCKET_MINIMUM) {
    stable_reorder(q, bucket_key);
    s->stats.bucketed_fronts++;
    s->stats.bucketed_particles += q->count;
  }
}

static void photon_front(shower_t* s) {
  bucket_front(s, &s->photons);
  queue_t in = s->photons;
  queue_t next;
  memset(&next, 0, sizeof next);
  uint64_t child = 0;
  uint64_t first = s->sched.next;
  for (uint64_t i = 0; i < in.count; ++i) {
    pout_t o;
    photon_source(s, &in.items[i], &o, child, first);
    for (uint32_t k = 0; k < o.next_count; ++k) push(&next, &o.next[k], CLS_REMAINING);
    for (uint32_t k = 0; k < o.charged_count; ++k) push(&s->leptons, &o.charged[k], CLS_PENDING);
    child += o.child_count;
  }
  s->sched.next = first + child;
  s->stats.photon_fronts++;
  if (in.count > s->stats.max_front) s->stats.max_front = in.count;
  queue_free(&in);
  s->photons = next;
}

static void lepton_front(shower_t* s) {
  bucket_front(s, &s->leptons);
  queue_t in = s->leptons;
  queue_t next;
  memset(&next, 0, sizeof next);
  uint64_t child = 0;
  uint64_t first = s->sched.next;
  for (uint64_t i = 0; i < in.count; ++i) {
    lout_t o;
    lepton_source(s, &in.items[i], &o, child, first);
    for (uint32_t k = 0; k < o.next_count; ++k) push(&next, &o.next[k], CLS_REMAINING);
    for (uint32_t k = 0; k < o.photon_count; ++k) push(&s->photons, &o.photons[k], CLS_PENDING);
    child += o.child_count;
  }
  s->sched.next = first + child;
  s->stats.lepton_fronts++;
  if (in.count > s->stats.max_front) s->stats.max_front = in.count;
  queue_free(&in);
  s->leptons = next;
}

/* The lepton reservation's multiplier and chunk (64 and 2^20 in the
 * reference); tests lower them to force history-range checkpoints. */
static uint64_t g_lepton_multiplier = 64u;
static uint64_t g_lepton_chunk = (uint64_t)1 << 20;

void oracle_em_set_lepton_reservation(uint64_t multiplier, uint64_t chunk) {
  g_lepton_multiplier = multiplier ? multiplier : 64u;
  g_lepton_chunk = chunk ? chunk : ((uint64_t)1 << 20);
}

/* History reservation of a call of n input particles: photons 2 n per front of
 * the call (reservePhotonSecondaryRange, EmRouter.inl:640-656); leptons
 * max(3 n, min(2^20, 64 * 3 n)) with its exclusive limit
 * (reserveResidentSecondaryRange, EmRouter.inl:658-679). */
static void open_call(shower_t* s, int lepton, uint64_t n) {
  sched_t* c = &s->sched;
  uint64_t count;
  if (!lepton) {
    count = n * 2u * c->photon_call_fronts;
  } else {
    uint64_t const minimum = 3u * n;
    count = minimum;
    if (minimum <= UINT64_MAX / c->lepton_multiplier) {
      count = minimum * c->lepton_multiplier < c->lepton_chunk ? minimum * c->lepton_multiplier : c->lepton_chunk;
      if (count < minimum) count = minimum;
    }
  }
  c->next = c->counter;
  c->limit = c->counter + count;
  c->counter += count;
  c->fronts = 0;
  c->open = 1;
  s->stats.reserved_history_ids += count;
  if (lepton) s->stats.lepton_calls++;
  else s->stats.photon_calls++;
}

/* Whether the shower's next front of `lepton` runs now: opens a call when the
 * shower enters that phase with particles (input ordered pending, remaining,
 * host), ends the call when its queue is empty, after the call's last front,
 * below the minimum batch after its first front, or (leptons) when the front
 * could exhaust the reserved range; an ended call passes to the other species. */
static int schedule_front(shower_t* s, int lepton) {
  sched_t* c = &s->sched;
  queue_t* q = lepton ? &s->leptons : &s->photons;
  if (c->phase != lepton) return 0;
  if (!c->open) {
    if (q->count == 0) {
      c->phase ^= 1;
      return 0;
    }
    stable_reorder(q, class_key);
    open_call(s, lepton, q->count);
  }
  uint64_t const n = q->count;
  int stop = n == 0 || c->fronts >= (lepton ? c->lepton_call_fronts : c->photon_call_fronts);
  if (!stop && c->fronts != 0 && n < c->min_batch) {
    stop = 1;
    s->stats.below_minimum_batch_checkpoints++;
  }
  if (!stop && lepton && n > (c->limit - c->next) / 3u) {
    stop = 1;
    s->stats.history_range_checkpoints++;
  }
  if (stop) {
    c->open = 0;
    c->phase ^= 1;
    for (uint64_t i = 0; i < q->count; ++i) q->cls[i] = CLS_REMAINING;
    return 0;
  }
  c->fronts++;
  return 1;
}

static int run_shower(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                      const c8mi_environment* environment, uint64_t shower_id,
                      uint64_t first_history_id, const c8mi_particle_state* primaries,
                      uint64_t primary_count, uint64_t maximum_fronts, int64_t* photons,
                      int64_t* electrons, int64_t* positrons, int64_t* energy_loss,
                      int64_t* counters, c8mi_particle_state* observations,
                      c8mi_em_fallback* fallbacks, or_em_stats* stats, double** tracks,
                      uint64_t* track_count, or_vertex** vertices, uint64_t* vertex_count,
                      c8mi_em_first_interaction* first, int* first_captured) {
  shower_t s;
  memset(&s, 0, sizeof s);
  s.record_tracks = tracks != NULL;
  s.record_vertices = vertices != NULL;
  s.x.t = tables;
  s.x.cont.t = tables;
  s.x.cont.em_cut_mev = config->em_transport_cut_mev;
  s.x.cont.muon_cut_mev = config->muon_transport_cut_mev;
  s.x.env = environment;
  s.x.cfg = config;
  s.x.seed = config->random_seed;
  s.x.shower = shower_id;
  s.sched.counter = first_history_id;
  s.sched.min_batch = config->resident_minimum_batch;
  s.sched.photon_call_fronts = PHOTON_CALL_FRONTS;
  s.sched.lepton_call_fronts = LEPTON_CALL_FRONTS;
  s.sched.lepton_multiplier = g_lepton_multiplier;
  s.sched.lepton_chunk = g_lepton_chunk;
  s.prof.photons = photons;
  s.prof.electrons = electrons;
  s.prof.positrons = positrons;
  s.prof.energy_loss = energy_loss;
  s.observations = observations;
  s.fallbacks = fallbacks;
  for (uint64_t i = 0; i < primary_count; ++i) {
    const ps_t* p = &primaries[i];
    ledger_add(&s.prof, config, OR_EMC_INJECTED, p->energy_gev * p->weight);
    if (p->pid == 22) push(&s.photons, p, CLS_HOST);
    else if (is_ee(p->pid)) push(&s.leptons, p, CLS_HOST);
    else {
      fb_t f = {1, FR_UNSUPPORTED_PARTICLE, *p, 0};
      record_fallback(&s, &f, STAGE_SELECTION);
    }
  }
  uint64_t rounds = 0;
  /* rounds of the device pool: the shower's photon front (if its schedule
   * runs one now), then its lepton front */
  while ((s.photons.count || s.leptons.count) && rounds < maximum_fronts) {
    if (schedule_front(&s, 0)) photon_front(&s);
    if (schedule_front(&s, 1)) lepton_front(&s);
    ++rounds;
  }
  double pending = 0.;
  for (uint64_t i = 0; i < s.photons.count; ++i)
    pending += s.photons.items[i].energy_gev * s.photons.items[i].weight;
  for (uint64_t i = 0; i < s.leptons.count; ++i)
    pending += s.leptons.items[i].energy_gev * s.leptons.items[i].weight;
  ledger_add(&s.prof, config, OR_EMC_PENDING, pending);
  s.stats.pending_photons = s.photons.count;
  s.stats.pending_leptons = s.leptons.count;
  s.stats.next_history_id = s.sched.counter;
  s.stats.observations = s.observation_count;
  memcpy(counters, s.prof.counters, sizeof s.prof.counters);
  *stats = s.stats;
  if (first && first_captured) {
    *first_captured = s.first_candidates != 0;
    if (*first_captured) *first = s.first;
  }
  queue_free(&s.photons);
  queue_free(&s.leptons);
  if (tracks) {
    if (s.track_overflow) {
      free(s.tracks);
      free(s.vertices);
      return -1;
    }
    *tracks = s.tracks;
    *track_count = s.track_count;
  }
  if (vertices) {
    if (s.vertex_overflow) {
      free(s.vertices);
      return -1;
    }
    *vertices = s.vertices;
    *vertex_count = s.vertex_count;
  }
  return 0;
}

int oracle_em_run_shower(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                         const c8mi_environment* environment, uint64_t shower_id,
                         uint64_t first_history_id, const c8mi_particle_state* primaries,
                         uint64_t primary_count, uint64_t maximum_fronts, int64_t* photons,
                         int64_t* electrons, int64_t* positrons, int64_t* energy_loss,
                         int64_t* counters, c8mi_particle_state* observations,
                         c8mi_em_fallback* fallbacks, or_em_stats* stats) {
  return run_shower(config, tables, environment, shower_id, first_history_id, primaries,
                    primary_count, maximum_fronts, photons, electrons, positrons, energy_loss,
                    counters, observations, fallbacks, stats, NULL, NULL, NULL, NULL, NULL, NULL);
}

int oracle_em_run_shower_tracks(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                                const c8mi_environment* environment, uint64_t shower_id,
                                uint64_t first_history_id, const c8mi_particle_state* primaries,
                                uint64_t primary_count, uint64_t maximum_fronts, int64_t* photons,
                                int64_t* electrons, int64_t* positrons, int64_t* energy_loss,
                                int64_t* counters, c8mi_particle_state* observations,
                                c8mi_em_fallback* fallbacks, or_em_stats* stats, double** tracks,
                                uint64_t* track_count) {
  if (!tracks || !track_count) return -1;
  *tracks = NULL;
  *track_count = 0;
  return run_shower(config, tables, environment, shower_id, first_history_id, primaries,
                    primary_count, maximum_fronts, photons, electrons, positrons, energy_loss,
                    counters, observations, fallbacks, stats, tracks, track_count, NULL, NULL, NULL, NULL);
}

void oracle_free(void* p) { free(p); }

int oracle_em_first_interaction(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                                const c8mi_environment* environment, uint64_t shower_id,
                                uint64_t first_history_id, const c8mi_particle_state* primaries,
                                uint64_t primary_count, c8mi_em_first_interaction* out, int* captured,
                                or_em_stats* stats) {
  uint64_t const bins = config->profile_bins ? config->profile_bins : 1;
  int64_t* hist = calloc(4 * bins, sizeof(int64_t));
  int64_t counters[OR_EM_COUNTERS];
  if (!hist || !out || !captured) {
    free(hist);
    return -1;
  }
  int rc = run_shower(config, tables, environment, shower_id, first_history_id, primaries, primary_count,
                      UINT64_MAX, hist, hist + bins, hist + 2 * bins, hist + 3 * bins, counters, NULL, NULL,
                      stats, NULL, NULL, NULL, NULL, out, captured);
  free(hist);
  return rc;
}

int oracle_em_run_shower_vertices(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                                  const c8mi_environment* environment, uint64_t shower_id,
                                  uint64_t first_history_id, const c8mi_particle_state* primaries,
                                  uint64_t primary_count, or_vertex** vertices, uint64_t* vertex_count,
                                  or_em_stats* stats) {
  if (!vertices || !vertex_count || !stats) return -1;
  *vertices = NULL;
  *vertex_count = 0;
  uint64_t const bins = config->profile_bins ? config->profile_bins : 1;
  int64_t* hist = calloc(4 * bins, sizeof(int64_t));
  int64_t counters[OR_EM_COUNTERS];
  if (!hist) return -1;
  int rc = run_shower(config, tables, environment, shower_id, first_history_id, primaries, primary_count,
                      UINT64_MAX, hist, hist + bins, hist + 2 * bins, hist + 3 * bins, counters, NULL, NULL,
                      stats, NULL, NULL, vertices, vertex_count, NULL, NULL);
  free(hist);
  return rc;
}

int oracle_em_select(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                     const c8mi_environment* environment, uint64_t shower_id,
                     const c8mi_particle_state* particle, or_selection* out) {
  ctx_t x;
  memset(&x, 0, sizeof x);
  x.t = tables;
  x.cont.t = tables;
  x.cont.em_cut_mev = config->em_transport_cut_mev;
  x.cont.muon_cut_mev = config->muon_transport_cut_mev;
  x.env = environment;
  x.cfg = config;
  x.seed = config->random_seed;
  x.shower = shower_id;
  inter_t in;
  fb_t fb = {0, 0, {0}, 0};
  select_interaction(&x, particle, &in, &fb);
  memset(out, 0, sizeof *out);
  out->status = in.status;
  out->process = in.process_id;
  out->component = in.component_hash;
  out->interaction_grammage = in.interaction_grammage;
  out->total_rate = in.total_rate;
  out->energy_fraction = in.energy_fraction;
  out->loss_quantile = in.loss_quantile;
  out->process_uniform = in.process_uniform;
  out->fallback = fb.flag;
  out->fallback_reason = fb.reason;
  return 0;
}

/* Independent showers in parallel over host threads (the CPU baseline of the
 * benchmark): shower k runs primaries[k] alone with shower_id first_shower + k,
 * into thread-private profile and record buffers; only the statistics are
 * returned. Each shower is the serial oracle_em_run_shower above. */
int oracle_em_run_batch(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                        const c8mi_environment* environment, uint64_t first_shower,
                        const c8mi_particle_state* primaries, uint64_t shower_count,
                        uint64_t maximum_fronts, int threads, or_em_stats* stats) {
  int failed = 0;
  uint64_t const bins = config->profile_bins ? config->profile_bins : 1;
#pragma omp parallel num_threads(threads > 0 ? threads : 1) reduction(| : failed)
  {
    int64_t* hist = calloc(4 * bins, sizeof(int64_t));
    int64_t counters[OR_EM_COUNTERS];
    c8mi_particle_state* obs = config->observation_capacity
                                   ? malloc(config->observation_capacity * sizeof *obs) : NULL;
    c8mi_em_fallback* fb = config->fallback_capacity
                               ? malloc(config->fallback_capacity * sizeof *fb) : NULL;
    if (!hist || (config->observation_capacity && !obs) || (config->fallback_capacity && !fb)) {
      failed = 1;
    } else {
#pragma omp for schedule(dynamic, 1)
      for (uint64_t k = 0; k < shower_count; ++k) {
        memset(hist, 0, 4 * bins * sizeof(int64_t));
        failed |= oracle_em_run_shower(config, tables, environment, first_shower + k, 2,
                                       &primaries[k], 1, maximum_fronts, hist, hist + bins,
                                       hist + 2 * bins, hist + 3 * bins, counters, obs, fb,
                                       &stats[k]) != 0;
      }
    }
    free(hist);
    free(obs);
    free(fb);
  }
  return failed ? -1 : 0;
}

/* oracle_em_run_batch that also keeps every shower's four histograms,
 * hist[k * 4 * profile_bins + h * profile_bins + b] (h: photons, electrons,
 * positrons, energy loss), and its counters[k * OR_EM_COUNTERS + c]. */
int oracle_em_run_batch_profiles(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                                 const c8mi_environment* environment, uint64_t first_shower,
                                 const c8mi_particle_state* primaries, uint64_t shower_count, int threads,
                                 or_em_stats* stats, int64_t* hist, int64_t* counters) {
  int failed = 0;
  uint64_t const bins = config->profile_bins ? config->profile_bins : 1;
#pragma omp parallel num_threads(threads > 0 ? threads : 1) reduction(| : failed)
  {
    c8mi_particle_state* obs = config->observation_capacity
                                   ? malloc(config->observation_capacity * sizeof *obs) : NULL;
    c8mi_em_fallback* fb = config->fallback_capacity
                               ? malloc(config->fallback_capacity * sizeof *fb) : NULL;
    if ((config->observation_capacity && !obs) || (config->fallback_capacity && !fb)) {
      failed = 1;
    } else {
#pragma omp for schedule(dynamic, 1)
      for (uint64_t k = 0; k < shower_count; ++k) {
        int64_t* h = hist + k * 4 * bins;
        memset(h, 0, 4 * bins * sizeof(int64_t));
        failed |= oracle_em_run_shower(config, tables, environment, first_shower + k, 2, &primaries[k], 1,
                                       UINT64_MAX, h, h + bins, h + 2 * bins, h + 3 * bins,
                                       counters + k * OR_EM_COUNTERS, obs, fb, &stats[k]) != 0;
      }
    }
    free(obs);
    free(fb);
  }
  return failed ? -1 : 0;
}

/* ===================== host handoff (CPU prefix + device calls) ===================== */

/* One side's accumulators (host steps or device fronts), swapped into the
 * shower around host steps. */
typedef struct {
  prof_t prof;
  ps_t* observations;
  uint64_t observation_count;
  c8mi_em_fallback* fallbacks;
  uint64_t fallback_count;
  double* tracks;
  uint64_t track_count, track_capacity;
  int track_overflow;
  c8mi_em_first_interaction first;
  uint64_t first_candidates;
} side_t;

static void swap_side(shower_t* s, side_t* d) {
  side_t cur = {s->prof, s->observations, s->observation_count, s->fallbacks, s->fallback_count,
                s->tracks, s->track_count, s->track_capacity, s->track_overflow, s->first, s->first_candidates};
  s->prof = d->prof;
  s->observations = d->observations;
  s->observation_count = d->observation_count;
  s->fallbacks = d->fallbacks;
  s->fallback_count = d->fallback_count;
  s->tracks = d->tracks;
  s->track_count = d->track_count;
  s->track_capacity = d->track_capacity;
  s->track_overflow = d->track_overflow;
  s->first = d->first;
  s->first_candidates = d->first_candidates;
  *d = cur;
}

/* HybridCascade::run (HybridCascade.inl:55-175) with EmRouter
 * (EmRouter.inl:64-423) for an EM shower: the host stack is LIFO
 * (CpuParticleScheduler::acquireNext, CpuParticleScheduler.inl:27-52); a
 * particle the router can route is staged, any other (and one the router
 * returned for a host step, marked cls = 1) takes one host step whose outputs
 * go on the stack in emission order. The router hands the staged batch to the
 * device when it holds at least the minimum batch, when device pending queues
 * are not empty, or after the maximum number of consecutive expansion rounds
 * (EmRouter.inl:246-423); a smaller batch is put back on the stack for one
 * host step each (scalar expansion, EmRouter.inl:260-293). */
typedef struct {
  shower_t s;      /* device side: queues, schedule, accumulators */
  side_t host;     /* host side (swapped in for host steps) */
  queue_t stack;   /* cls 1: returned by an expansion, steps on the host */
  queue_t staged;
  queue_t expanded; /* states of the last expansion round */
  int force_small;
  uint64_t consecutive;
  uint64_t min_batch, max_rounds;
  c8mi_em_limits limits;
  c8mi_em_hybrid_result r;
  int prefix_only, stop;
  uint64_t fallback_cursor;       /* device fallback records already returned or kept */
  c8mi_em_config without_moliere; /* the configuration of a fallback's host step */
} hybrid_t;

static int host_can_route(const hybrid_t* h, const ps_t* p) {
  /* EmRouter::canRoute (EmRouter.inl:64-105): an EM particle the backend can
   * transport (EmBackend canTransport: finite state, positive energy,
   * non-negative weight) inside the environment. */
  if (!(p->pid == 22 || is_ee(p->pid))) return 0;
  if (!finite(p->energy_gev) || !(p->energy_gev > 0.) || !finite(p->weight) || p->weight < 0.) return 0;
  for (int a = 0; a < 3; ++a)
    if (!finite(p->position_m[a]) || !finite(p->direction[a])) return 0;
  layer_t l = query_environment(h->s.x.env, p->position_m, p->direction, p->environment_node);
  return l.status == ES_SUCCESS;
}

static int same_transport_state(const ps_t* a, const ps_t* b) {
  if (a->pid != b->pid || a->environment_node != b->environment_node || a->medium_id != b->medium_id ||
      a->energy_gev != b->energy_gev || a->time_s != b->time_s || a->weight != b->weight)
    return 0;
  for (int k = 0; k < 3; ++k)
    if (a->position_m[k] != b->position_m[k] || a->direction[k] != b->direction[k]) return 0;
  return 1;
}

static void host_stage(hybrid_t* h, const ps_t* p) {
  /* EmRouter::stage (EmRouter.inl:115-137) */
  for (uint64_t i = 0; i < h->expanded.count; ++i)
    if (h->expanded.items[i].history_id == p->history_id && same_transport_state(&h->expanded.items[i], p)) {
      h->force_small = 1;
      break;
    }
  push(&h->staged, p, 0);
  h->r.particles_staged++;
}

/* One host step: outputs with history IDs from the shower's counter, a
 * continuation alone or the children in ID order. */
static uint32_t host_step(hybrid_t* h, const ps_t* p, ps_t out[4], int moliere) {
  shower_t* s = &h->s;
  const c8mi_em_config* const cfg = s->x.cfg;
  if (!moliere) s->x.cfg = &h->without_moliere;
  or_em_stats const before = s->stats;
  uint64_t const first = s->sched.counter;
  uint32_t n = 0, children;
  swap_side(s, &h->host);
  uint64_t const tracks_before = s->track_count;
  if (p->pid == 22) {
    pout_t o;
    photon_source(s, p, &o, 0, first);
    for (uint32_t k = 0; k < o.next_count; ++k) out[n++] = o.next[k];
    for (uint32_t k = 0; k < o.charged_count; ++k) out[n++] = o.charged[k];
    children = o.child_count;
  } else {
    lout_t o;
    lepton_source(s, p, &o, 0, first);
    for (uint32_t k = 0; k < o.next_count; ++k) out[n++] = o.next[k];
    for (uint32_t k = 0; k < o.photon_count; ++k) out[n++] = o.photons[k];
    children = o.child_count;
  }
  h->r.cpu_tracks += s->track_count - tracks_before;
  swap_side(s, &h->host);
  s->x.cfg = cfg;
  if (children && n == 2 && out[0].history_id > out[1].history_id) {
    ps_t t = out[0];
    out[0] = out[1];
    out[1] = t;
  }
  s->sched.counter += children;
  h->r.cpu_particles++;
  h->r.cpu_photon_steps += s->stats.photon_steps - before.photon_steps;
  h->r.cpu_lepton_steps += s->stats.lepton_steps - before.lepton_steps;
  h->r.cpu_interactions += (s->stats.photon_interactions - before.photon_interactions) +
                           (s->stats.lepton_interactions - before.lepton_interactions);
  h->r.cpu_secondaries += s->stats.secondaries - before.secondaries;
  s->stats = before;
  return n;
}

static int history_step_less(const void* a, const void* b) {
  const ps_t* x = a;
  const ps_t* y = b;
  if (x->history_id != y->history_id) return x->history_id < y->history_id ? -1 : 1;
  return x->step_id < y->step_id ? -1 : x->step_id > y->step_id;
}

/* EmRouter::returnFallbacks (EmRouter.inl:709-800): the device fallbacks of the
 * call; a Molière sampling failure goes back to the host stack for a host step
 * without Molière scattering (marked cls 2), in (history, step) order; the
 * others stay records. */
static void return_fallbacks(hybrid_t* h) {
  shower_t* s = &h->s;
  uint64_t const total = s->fallback_count;
  if (total <= h->fallback_cursor) return;
  uint64_t const stored = total < s->x.cfg->fallback_capacity ? total : s->x.cfg->fallback_capacity;
  queue_t returned;
  memset(&returned, 0, sizeof returned);
  for (uint64_t i = h->fallback_cursor; i < stored; ++i) {
    const c8mi_em_fallback* f = &s->fallbacks[i];
    if (f->reason == FR_MOLIERE_FAILED) push(&returned, &f->particle, 0);
    else h->r.fallbacks_kept++;
  }
  uint64_t const seen = stored > h->fallback_cursor ? stored : h->fallback_cursor;
  if (total > seen) h->r.fallbacks_kept += total - seen;
  if (returned.count) qsort(returned.items, returned.count, sizeof(ps_t), history_step_less);
  for (uint64_t i = 0; i < returned.count; ++i) {
    const ps_t* p = &returned.items[i];
    push(&h->stack, p, 2);
    ledger_add(&h->host.prof, s->x.cfg, OR_EMC_INJECTED, p->energy_gev * p->weight);
    h->r.fallbacks_returned++;
  }
  queue_free(&returned);
  h->fallback_cursor = total;
}

/* A wavefront call (IEmBackend::run{Photon,Lepton}Wavefront): the device
 * pending prefix and the host particles, the call's fronts, its remaining
 * particles back to the staged batch. The energy the device keeps leaves the
 * host ledger for the device's. */
static void device_call(hybrid_t* h, int lepton, const queue_t* host) {
  shower_t* s = &h->s;
  const c8mi_em_config* cfg = s->x.cfg;
  queue_t* q = lepton ? &s->leptons : &s->photons;
  for (uint64_t i = 0; i < host->count; ++i) {
    const ps_t* p = &host->items[i];
    push(q, p, CLS_HOST);
    ledger_add(&s->prof, cfg, OR_EMC_INJECTED, p->energy_gev * p->weight);
    ledger_add(&h->host.prof, cfg, OR_EMC_INJECTED, -(p->energy_gev * p->weight));
  }
  uint64_t const n = q->count;
  if (n > h->r.maximum_input_batch) h->r.maximum_input_batch = n;
  s->sched.phase = lepton;
  s->sched.open = 0;
  while (schedule_front(s, lepton)) {
    if (lepton) lepton_front(s);
    else photon_front(s);
  }
  for (uint64_t i = 0; i < q->count; ++i) {
    const ps_t* p = &q->items[i];
    push(&h->staged, p, 0);
    ledger_add(&s->prof, cfg, OR_EMC_INJECTED, -(p->energy_gev * p->weight));
    ledger_add(&h->host.prof, cfg, OR_EMC_INJECTED, p->energy_gev * p->weight);
  }
  q->count = 0;
  return_fallbacks(h);
}

static uint64_t device_pending(const hybrid_t* h) { return h->s.photons.count + h->s.leptons.count; }

/* EmRouter::advanceOneWavefrontAndReturn (EmRouter.inl:246-423). */
static void router_advance(hybrid_t* h) {
  uint64_t const pending = device_pending(h);
  if (h->staged.count == 0 && pending == 0) return;
  if (h->staged.count < h->min_batch && !h->force_small && pending == 0 && h->consecutive < h->max_rounds) {
    /* scalar expansion: the batch goes back to the stack for one host step each */
    h->expanded.count = 0;
    for (uint64_t i = 0; i < h->staged.count; ++i) {
      push(&h->stack, &h->staged.items[i], 1);
      push(&h->expanded, &h->staged.items[i], 0);
    }
    h->r.particles_returned_for_expansion += h->staged.count;
    h->staged.count = 0;
    h->force_small = 0;
    h->consecutive++;
    h->r.small_batch_expansions++;
    return;
  }
  if (h->staged.count < h->min_batch) {
    if (h->force_small) h->r.stalled_boundary_flushes++;
    else if (pending == 0 && h->consecutive >= h->max_rounds) h->r.scalar_expansion_budget_flushes++;
  }
  h->force_small = 0;
  h->consecutive = 0;
  h->expanded.count = 0;
  if (h->prefix_only) {
    h->stop = 1;
    return;
  }
  queue_t current = h->staged;
  memset(&h->staged, 0, sizeof h->staged);
  if (current.count > h->limits.maximum_resident_input_batch) {
    for (uint64_t i = h->limits.maximum_resident_input_batch; i < current.count; ++i)
      push(&h->staged, &current.items[i], 0);
    current.count = h->limits.maximum_resident_input_batch;
    h->r.input_batch_splits++;
  }
  h->r.wavefronts++;
  queue_t photons, leptons;
  memset(&photons, 0, sizeof photons);
  memset(&leptons, 0, sizeof leptons);
  for (uint64_t i = 0; i < current.count; ++i)
    push(current.items[i].pid == 22 ? &photons : &leptons, &current.items[i], 0);
  queue_free(&current);
  uint64_t const max_photons = h->limits.maximum_resident_photon_batch;
  uint64_t const pending_photons = h->s.photons.count < max_photons ? h->s.photons.count : max_photons;
  if (photons.count > max_photons - pending_photons) {
    for (uint64_t i = max_photons - pending_photons; i < photons.count; ++i) push(&h->staged, &photons.items[i], 0);
    photons.count = max_photons - pending_photons;
    h->r.input_batch_splits++;
  }
  if (photons.count + pending_photons) device_call(h, 0, &photons);
  uint64_t const max_leptons = h->limits.maximum_resident_lepton_batch;
  uint64_t const pending_leptons = h->s.leptons.count < max_leptons ? h->s.leptons.count : max_leptons;
  if (leptons.count + h->s.leptons.count) {
    if (leptons.count > max_leptons - pending_leptons) {
      for (uint64_t i = max_leptons - pending_leptons; i < leptons.count; ++i)
        push(&h->staged, &leptons.items[i], 0);
      leptons.count = max_leptons - pending_leptons;
      h->r.input_batch_splits++;
    }
    device_call(h, 1, &leptons);
  }
  queue_free(&photons);
  queue_free(&leptons);
}

static int router_ready(const hybrid_t* h) {
  /* EmRouter::readyForScalarInterleave (EmRouter.inl:150-167) */
  return h->force_small || device_pending(h) != 0 || h->staged.count >= h->min_batch;
}

static int router_pending(const hybrid_t* h) { return h->staged.count != 0 || device_pending(h) != 0; }

static void hybrid_cascade(hybrid_t* h) {
  while (!h->stop && (h->stack.count || router_pending(h))) {
    while (!h->stop && h->stack.count) {
      ps_t const p = h->stack.items[--h->stack.count];
      int const marked = h->stack.cls[h->stack.count];
      if (!marked && host_can_route(h, &p)) {
        host_stage(h, &p);
        while (!h->stop && router_ready(h)) router_advance(h);
        continue;
      }
      ps_t out[4];
      uint32_t const n = host_step(h, &p, out, marked != 2);
      for (uint32_t k = 0; k < n; ++k) push(&h->stack, &out[k], 0);
    }
    if (!h->stop && router_pending(h)) router_advance(h);
  }
}

int oracle_em_run_hybrid(const c8mi_em_config* config, const c8mi_native_table_set* tables,
                         const c8mi_environment* environment, uint64_t shower_id, uint64_t first_history_id,
                         const c8mi_particle_state* primaries, uint64_t primary_count,
                         const c8mi_em_hybrid_config* hc, const c8mi_em_limits* limits, int prefix_only,
                         or_hybrid_side* host, or_hybrid_side* device, c8mi_particle_state** staged,
                         uint64_t* staged_count, c8mi_em_hybrid_result* out) {
  hybrid_t* h = calloc(1, sizeof *h);
  if (!h) return -1;
  shower_t* s = &h->s;
  s->record_tracks = 1;
  s->x.t = tables;
  s->x.cont.t = tables;
  s->x.cont.em_cut_mev = config->em_transport_cut_mev;
  s->x.cont.muon_cut_mev = config->muon_transport_cut_mev;
  s->x.env = environment;
  s->x.cfg = config;
  s->x.seed = config->random_seed;
  s->x.shower = shower_id;
  s->sched.counter = first_history_id;
  h->min_batch = hc && hc->minimum_batch_size ? hc->minimum_batch_size : 16384u;
  h->max_rounds = hc && hc->maximum_expansion_rounds ? hc->maximum_expansion_rounds : 8u;
  s->sched.min_batch = h->min_batch;
  s->sched.photon_call_fronts = hc && hc->photon_wavefronts ? hc->photon_wavefronts : PHOTON_CALL_FRONTS;
  s->sched.lepton_call_fronts = hc && hc->lepton_wavefronts ? hc->lepton_wavefronts : LEPTON_CALL_FRONTS;
  s->sched.lepton_multiplier = hc && hc->lepton_reservation_multiplier ? hc->lepton_reservation_multiplier : 64u;
  s->sched.lepton_chunk = hc && hc->lepton_reservation_chunk ? hc->lepton_reservation_chunk : ((uint64_t)1 << 20);
  if (limits) {
    h->limits = *limits;
  } else {
    h->limits.maximum_resident_photon_batch = UINT64_MAX;
    h->limits.maximum_resident_lepton_batch = UINT64_MAX;
    h->limits.maximum_resident_input_batch = UINT64_MAX;
  }
  h->prefix_only = prefix_only;
  h->without_moliere = *config;
  h->without_moliere.moliere_enabled = 0;
  uint64_t const bins = config->profile_bins;
  s->prof.photons = device->hist;
  s->prof.electrons = device->hist + bins;
  s->prof.positrons = device->hist + 2 * bins;
  s->prof.energy_loss = device->hist + 3 * bins;
  s->observations = device->observations;
  s->fallbacks = device->fallbacks;
  h->host.prof.photons = host->hist;
  h->host.prof.electrons = host->hist + bins;
  h->host.prof.positrons = host->hist + 2 * bins;
  h->host.prof.energy_loss = host->hist + 3 * bins;
  h->host.observations = host->observations;
  h->host.fallbacks = host->fallbacks;
  for (uint64_t i = 0; i < primary_count; ++i) {
    ledger_add(&h->host.prof, config, OR_EMC_INJECTED, primaries[i].energy_gev * primaries[i].weight);
    push(&h->stack, &primaries[i], 0);
  }
  hybrid_cascade(h);
  int const stranded = prefix_only && h->stack.count != 0;
  if (prefix_only) {
    /* the handed batch leaves the host ledger */
    for (uint64_t i = 0; i < h->staged.count; ++i)
      ledger_add(&h->host.prof, config, OR_EMC_INJECTED,
                 -(h->staged.items[i].energy_gev * h->staged.items[i].weight));
  }
  c8mi_em_hybrid_result* r = &h->r;
  r->photon_calls = s->stats.photon_calls;
  r->lepton_calls = s->stats.lepton_calls;
  r->device_photon_fronts = s->stats.photon_fronts;
  r->device_lepton_fronts = s->stats.lepton_fronts;
  r->device_photon_steps = s->stats.photon_steps;
  r->device_lepton_steps = s->stats.lepton_steps;
  r->device_interactions = s->stats.photon_interactions + s->stats.lepton_interactions;
  r->device_secondaries = s->stats.secondaries;
  r->bucketed_fronts = s->stats.bucketed_fronts;
  r->reserved_history_ids = s->stats.reserved_history_ids;
  r->below_minimum_batch_checkpoints = s->stats.below_minimum_batch_checkpoints;
  r->history_range_checkpoints = s->stats.history_range_checkpoints;
  r->next_history_id = s->sched.counter;
  *out = *r;
  memcpy(device->counters, s->prof.counters, sizeof s->prof.counters);
  memcpy(host->counters, h->host.prof.counters, sizeof h->host.prof.counters);
  device->observation_count = s->observation_count;
  device->fallback_count = s->fallback_count;
  device->tracks = s->tracks;
  device->track_count = s->track_count;
  host->observation_count = h->host.observation_count;
  host->fallback_count = h->host.fallback_count;
  host->tracks = h->host.tracks;
  host->track_count = h->host.track_count;
  host->first = h->host.first;
  host->first_captured = h->host.first_candidates != 0;
  device->first = s->first;
  device->first_captured = s->first_candidates != 0;
  int const overflow = s->track_overflow || h->host.track_overflow;
  if (staged && staged_count) {
    *staged = h->staged.items;
    *staged_count = h->staged.count;
    h->staged.items = NULL;
  }
  queue_free(&h->staged);
  queue_free(&h->stack);
  queue_free(&h->expanded);
  queue_free(&s->photons);
  queue_free(&s->leptons);
  free(h);
  return overflow ? -1 : stranded ? -2 : 0;
}

/* ---- single-function probes for the tests ---- */

int oracle_em_environment_query(const c8mi_environment* e, const double position[3],
                                const double direction[3], int32_t node, int32_t* out_node,
                                double* out_density) {
  layer_t l = query_environment(e, position, direction, node);
  *out_node = l.node;
  *out_density = l.density;
  return l.status;
}

int oracle_em_boundary(const c8mi_environment* e, const double position[3],
                       const double direction[3], int32_t node, double* distance,
                       int32_t* next_node) {
  boundary_t b = distance_to_boundary(e, position, direction, node);
  *distance = b.distance;
  *next_node = b.next_node;
  return b.status;
}

int oracle_em_grammage(const c8mi_environment* e, int32_t node, const double position[3],
                       const double direction[3], double distance, double* out) {
  gq_t g = environment_grammage(e, node, position, direction, distance);
  *out = g.value;
  return g.status;
}

int oracle_em_arclength(const c8mi_environment* e, int32_t node, const double position[3],
                        const double direction[3], double grammage, double* out) {
  gq_t g = environment_arclength(e, node, position, direction, grammage);
  *out = g.value;
  return g.status;
}

int oracle_em_pair_lpm(const c8mi_photon_pair_lpm* s, uint64_t comp, double e_mev, double x,
                       double rho, double* out) {
  *out = 0.;
  return pair_lpm(s, comp, e_mev, x, rho, out);
}

int oracle_em_brems_lpm(const c8mi_brems_lpm* s, uint64_t comp, double e_mev, double v,
                        double rho, double* out) {
  *out = 0.;
  return brems_lpm(s, comp, e_mev, v, rho, out);
}

int oracle_em_thinning(const c8mi_em_thinning* c, double pe, double pw, double e1, double e2,
                       double u1, double u2, uint32_t* keep, double* w1, double* w2) {
  thin_t t = em_thinning(c, pe, pw, e1, e2, u1, u2);
  *keep = t.keep;
  *w1 = t.w1;
  *w2 = t.w2;
  return t.status;
}

int oracle_em_pair_split(double e_gev, double x, double* normalized, double* decoded) {
  if (!encode_pair_split(e_gev, x, normalized)) return 0;
  return decode_pair_split(e_gev, *normalized, decoded) ? 1 : 0;
}

double oracle_em_sauter_cosine(double e, double u) { return sauter_cosine(e, u); }

int oracle_em_compton_cosines(double e, double v, double* pc, double* ec) {
  return compton_cosines(e, v, pc, ec);
}

int oracle_em_annihilation_rho(double e, double m, double u, double* rho) {
  return annihilation_rho(e, m, u, rho);
}
