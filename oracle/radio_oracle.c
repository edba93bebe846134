/*
 * c8mi oracle — TEST INFRASTRUCTURE ONLY (see c8_oracle.h).
 *
 * FP64 restatement of the CoREAS / ZHS endpoint projection of CORSIKA8-kokkos,
 * corsika/detail/kokkos/radio/algorithms/RadioProjection.inl, in plain C.
 * Operation order follows the reference expression by expression so that the
 * int64 fixed-point waveforms are comparable bit for bit; compile without
 * floating-point contraction.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "c8_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* corsika/kokkos/ScalarPhysicalConstants.hpp:19-21 and
 * corsika/kokkos/radio/RadioProjectionData.hpp:21-37 */
static const double kC = 299792458.;
static const double kE = 1.6021766208e-19;
static const double kEps0 = 8.8541878128e-12;
static const double kPi = 3.141592653589793238462643383279502884;
#define OR_EM_CONSTANT (1. / (4. * kPi * kEps0 * kC))
static const double kCoreasThreshold = 1.e-3;
static const double kHeadroom = 4611686018427387904.; /* 0x1p62 */
static const double kIntLimit = 9223372036854775808.;  /* 0x1p63 */

typedef struct { double x, y, z; } vec3;

static vec3 v_add(vec3 a, vec3 b) { vec3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static vec3 v_sub(vec3 a, vec3 b) { vec3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static vec3 v_mul(vec3 a, double s) { vec3 r = {a.x * s, a.y * s, a.z * s}; return r; }
static vec3 v_div(vec3 a, double s) { vec3 r = {a.x / s, a.y / s, a.z / s}; return r; }
static double v_dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static vec3 v_cross(vec3 a, vec3 b) {
  vec3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  return r;
}
static double v_norm(vec3 a) { return sqrt(v_dot(a, a)); }

/* observer_nearest_bin, RadioProjection.inl:75-82: integer/fraction split so a
 * fraction just below one half never rounds up. */
static double nearest_bin(double scaled) {
  double integral = floor(scaled);
  return integral + (scaled - integral >= 0.5 ? 1. : 0.);
}

/* co_reas_doppler, RadioProjection.inl:113-143. The host build evaluates the
 * cancelling case in long double; the device build uses an error-free
 * two-product/two-sum expansion. Both are restated. */
static double doppler(double index, vec3 beta, vec3 emit, int device_variant) {
  double ordinary = 1. - index * v_dot(beta, emit);
  if (ordinary != 0.) return ordinary;
  if (!device_variant) {
    long double s = (long double)beta.x * emit.x + (long double)beta.y * emit.y +
                    (long double)beta.z * emit.z;
    return (double)(1.L - (long double)index * s);
  }
  double px = beta.x * emit.x, py = beta.y * emit.y, pz = beta.z * emit.z;
  double ex = fma(beta.x, emit.x, -px);
  double ey = fma(beta.y, emit.y, -py);
  double ez = fma(beta.z, emit.z, -pz);
  double xy = px + py;
  double xyv = xy - px;
  double xye = (px - (xy - xyv)) + (py - xyv);
  double xyz = xy + pz;
  double xyzv = xyz - xy;
  double xyze = (xy - (xyz - xyzv)) + (pz - xyzv);
  double low = ex + ey + ez + xye + xyze;
  double product = index * xyz;
  double perr = fma(index, xyz, -product) + index * low;
  return (1. - product) - perr;
}

typedef struct {
  vec3 start, end, displacement, beta;
  double t0, t1, duration, length, beta_module, constant;
} track_t;

/* make_radio_track_kinematics, RadioProjection.inl:145-168, from the compact
 * track: the charge sign travels in the weight's sign, so
 * constant = (e * EmConstant) * signed_weight equals the reference's
 * (charge * EmConstant) * weight bit for bit. */
static int make_track(const double* tracks, uint64_t stride, uint64_t i, track_t* t) {
  t->start.x = tracks[0 * stride + i];
  t->start.y = tracks[1 * stride + i];
  t->start.z = tracks[2 * stride + i];
  t->end.x = tracks[3 * stride + i];
  t->end.y = tracks[4 * stride + i];
  t->end.z = tracks[5 * stride + i];
  t->t0 = tracks[6 * stride + i];
  t->t1 = tracks[7 * stride + i];
  t->duration = t->t1 - t->t0;
  t->displacement = v_sub(t->end, t->start);
  t->length = v_norm(t->displacement);
  if (!(t->duration > 0.) || !(t->length > 0.)) return 0;
  t->beta = v_div(t->displacement, kC * t->duration);
  t->beta_module = t->length / (kC * t->duration);
  t->constant = kE * OR_EM_CONSTANT * tracks[8 * stride + i];
  return 1;
}

typedef struct {
  double time, n_source, n_destination, distance;
  vec3 emit;
} path_t;

/* propagate, RadioProjection.inl:170-238 (straight ray, nearest-bin table with
 * the released edge extrapolation and its shadowed-index quirks). */
static path_t propagate(const c8mi_radio_propagation* t, vec3 source, vec3 destination) {
  path_t p;
  vec3 displacement = v_sub(destination, source);
  double distance = v_norm(displacement);
  vec3 emit = v_div(displacement, distance);
  if (t->homogeneous_refractive_index > 0.) {
    double index = t->homogeneous_refractive_index;
    p.time = index * distance / kC;
    p.n_source = index;
    p.n_destination = index;
    p.distance = distance;
    p.emit = emit;
    return p;
  }
  double source_height = (source.z - t->minimum_height_m) * t->inverse_step_per_m;
  double destination_height = (destination.z - t->minimum_height_m) * t->inverse_step_per_m;
  uint64_t last = t->table_size - 1;
  uint64_t destination_index = (uint64_t)(destination_height + 0.5);
  double n_source = 1.;
  double integrated_source = 1.;
  double n_destination = t->refractivity[destination_index] + 1.;
  double integrated_destination = t->integrated_refractivity[destination_index];
  double height = 1.;
  if (source_height + 0.5 >= (double)last) {
    n_source = t->refractivity[last] +
               t->slope_refractivity_upper *
                   fabs((source.z - t->maximum_height_m) * t->inverse_step_per_m) +
               1.;
    integrated_source = t->integrated_refractivity[last] +
                        t->slope_integrated_refractivity_upper *
                            fabs(t->maximum_height_m - source.z);
    height = source_height - destination_height;
  } else if (source_height + 0.5 < (double)last && source_height > 0.) {
    uint64_t source_index = (uint64_t)(source_height + 0.5);
    n_source = t->refractivity[source_index] + 1.;
    integrated_source = t->integrated_refractivity[source_index];
    height = (double)((int64_t)source_index - (int64_t)destination_index) * t->step_m;
    if (height == 0.) height = 1.;
  } else if (source_height == 0.) {
    n_source = 1.;
    n_destination = 1.;
    integrated_source = t->integrated_refractivity[0];
    height = destination_height - source_height;
  } else {
    n_source = t->refractivity[0] + t->slope_refractivity_lower * fabs(source_height) + 1.;
    integrated_source = t->integrated_refractivity[0] +
                        t->slope_integrated_refractivity_lower * fabs(source_height);
    n_destination = 1.;
    height = destination_height - fabs(source_height);
  }
  p.time = (1. + (integrated_source - integrated_destination) / height) * distance / kC;
  p.n_source = n_source;
  p.n_destination = n_destination;
  p.distance = distance;
  p.emit = emit;
  return p;
}

typedef struct {
  double position[3], start_time, duration, sample_rate, fixed_scale;
  uint64_t nbins, offset;
} obs_t;

typedef struct {
  int64_t* fixed; /* 3 blocks of total_bins, or NULL */
  double* floating;
  uint64_t total_bins;
} wave_t;

/* checked_radio_atomic_add + add_radio_fixed_point, RadioProjection.inl:245-260.
 * The increment is always applied (two's complement wrap, like the atomic);
 * the return value reports whether the running sum overflowed. */
static int add_fixed(int64_t* address, double contribution, double scale) {
  double scaled = contribution * scale;
  if (!isfinite(scaled) || fabs(scaled) >= kIntLimit) return 0;
  int64_t increment = (int64_t)rint(scaled);
  int64_t previous = *address;
  *address = (int64_t)((uint64_t)previous + (uint64_t)increment);
  return !((increment > 0 && previous > INT64_MAX - increment) ||
           (increment < 0 && previous < INT64_MIN - increment));
}

/* add_sample, RadioProjection.inl:262-301. CoREAS fields are gated to the
 * observer window; ZHS potentials are only bin-range checked. */
static void add_sample(const obs_t* o, wave_t* w, double time, vec3 c, int potential,
                       uint64_t* contribution_counter, or_radio_counters* counters) {
  if (!potential) {
    if (time < o->start_time || time > o->start_time + o->duration) return;
  }
  double bin_value = nearest_bin((time - o->start_time) * o->sample_rate);
  if (!(bin_value >= 0. && bin_value < (double)o->nbins)) return;
  uint64_t bin = (uint64_t)bin_value;
  if (bin >= o->nbins) return;
  uint64_t offset = o->offset + bin;
  if (w->floating) {
    w->floating[offset] += c.x;
    w->floating[w->total_bins + offset] += c.y;
    w->floating[2 * w->total_bins + offset] += c.z;
  }
  if (w->fixed) {
    int vx = add_fixed(w->fixed + offset, c.x, o->fixed_scale);
    int vy = add_fixed(w->fixed + w->total_bins + offset, c.y, o->fixed_scale);
    int vz = add_fixed(w->fixed + 2 * w->total_bins + offset, c.z, o->fixed_scale);
    if (!(vx && vy && vz)) {
      counters->fixed_point_overflows += 1;
      return;
    }
  }
  *contribution_counter += 1;
}

/* separate_same_bin_endpoints, RadioProjection.inl:303-336. */
static void separate_same_bin(double* start_time, double* end_time, double grid,
                              int signed_order) {
  long long start_bin = (long long)nearest_bin(*start_time / grid);
  long long end_bin = (long long)nearest_bin(*end_time / grid);
  if (start_bin != end_bin) return;
  double sf = *start_time / grid - floor(*start_time / grid);
  double ef = *end_time / grid - floor(*end_time / grid);
  int forward = !signed_order || *end_time >= *start_time;
  if (forward) {
    if (sf >= 0.5 && ef >= 0.5)
      *start_time -= grid;
    else if (sf < 0.5 && ef < 0.5)
      *end_time += grid;
    else if (ef >= 1. - sf)
      *end_time += grid;
    else
      *start_time -= grid;
  } else {
    if (sf >= 0.5 && ef >= 0.5)
      *end_time -= grid;
    else if (sf < 0.5 && ef < 0.5)
      *start_time += grid;
    else if (sf >= 1. - ef)
      *start_time += grid;
    else
      *end_time -= grid;
  }
}

static vec3 transverse(vec3 emit, vec3 beta) { return v_cross(emit, v_cross(emit, beta)); }

/* accumulate_co_reas, RadioProjection.inl:338-409. */
static void coreas_pair(const track_t* t, const c8mi_radio_propagation* prop, const obs_t* o,
                        wave_t* w, or_radio_counters* counters) {
  vec3 dest = {o->position[0], o->position[1], o->position[2]};
  path_t ps = propagate(prop, t->start, dest);
  path_t pe = propagate(prop, t->end, dest);
  double pre = doppler(ps.n_source, t->beta, ps.emit, 0);
  double post = doppler(pe.n_source, t->beta, pe.emit, 0);
  double start_receive = t->t0 + ps.time;
  double end_receive = t->t1 + pe.time;
  if (ps.n_destination > 1. &&
      (fabs(pre) < kCoreasThreshold || fabs(post) < kCoreasThreshold)) {
    vec3 midpoint = v_mul(v_add(t->start, t->end), 0.5);
    double middle_time = (t->t0 + t->t1) * 0.5;
    path_t pm = propagate(prop, midpoint, dest);
    double middle_receive = middle_time + pm.time;
    double md = doppler(pm.n_source, t->beta, pm.emit, 0);
    vec3 fs = v_mul(transverse(pm.emit, t->beta),
                    t->constant * o->sample_rate / (md * pm.distance));
    vec3 fe = v_mul(fs, -1.);
    double delta = t->length / (kC * t->beta_module) * fabs(md);
    if (start_receive < end_receive) {
      start_receive = middle_receive - 0.5 * delta;
      end_receive = middle_receive + 0.5 * delta;
    } else {
      start_receive = middle_receive + 0.5 * delta;
      end_receive = middle_receive - 0.5 * delta;
    }
    double grid = 1. / o->sample_rate;
    delta = end_receive - start_receive;
    if (fabs(delta) < grid) {
      double scale = fabs(delta / grid);
      fs = v_mul(fs, scale);
      fe = v_mul(fe, scale);
      separate_same_bin(&start_receive, &end_receive, grid, 1);
    }
    add_sample(o, w, start_receive, fs, 0, &counters->coreas_contributions, counters);
    add_sample(o, w, end_receive, fe, 0, &counters->coreas_contributions, counters);
    return;
  }
  vec3 fs = v_mul(transverse(ps.emit, t->beta),
                  t->constant * o->sample_rate / (pre * ps.distance));
  vec3 fe = v_mul(transverse(pe.emit, t->beta),
                  -t->constant * o->sample_rate / (post * pe.distance));
  if (pre < 1.e-9 || post < 1.e-9) {
    double grid = 1. / o->sample_rate;
    double delta = end_receive - start_receive;
    if (fabs(delta) < grid) {
      double scale = fabs(delta / grid);
      fs = v_mul(fs, scale);
      fe = v_mul(fe, scale);
      separate_same_bin(&start_receive, &end_receive, grid, 0);
    }
  }
  add_sample(o, w, start_receive, fs, 0, &counters->coreas_contributions, counters);
  add_sample(o, w, end_receive, fe, 0, &counters->coreas_contributions, counters);
}

/* accumulate_zhs_segment, RadioProjection.inl:411-480. */
static void zhs_segment(vec3 p1, vec3 p2, double time1, double time2, vec3 beta,
                        double constant, const c8mi_radio_propagation* prop, const obs_t* o,
                        wave_t* w, or_radio_counters* counters, int subdivided) {
  vec3 dest = {o->position[0], o->position[1], o->position[2]};
  vec3 midpoint = v_mul(v_add(p1, p2), 0.5);
  path_t path = propagate(prop, midpoint, dest);
  double n_source = path.n_source;
  double bk = v_dot(beta, path.emit);
  double middle_time = (time1 + time2) * 0.5;
  double d1 = time1 + path.time - n_source * bk * (time1 - middle_time);
  double d2 = time2 + path.time - n_source * bk * (time2 - middle_time);
  double sign = 1.;
  if (d1 > d2) {
    d1 = time2 + path.time - n_source * bk * (time2 - middle_time);
    d2 = time1 + path.time - n_source * bk * (time1 - middle_time);
    sign = -1.;
  }
  double start_bin = floor((d1 - o->start_time) * o->sample_rate + 0.5);
  double end_bin = floor((d2 - o->start_time) * o->sample_rate + 0.5);
  vec3 bperp = v_cross(path.emit, v_cross(beta, path.emit));
  double den = 1. - n_source * bk;
  if (start_bin == end_bin) {
    vec3 potential;
    if (fabs(den) > 1.e-15) {
      double fraction = fabs(d2 * o->sample_rate - d1 * o->sample_rate);
      potential = v_mul(bperp, sign * constant * fraction / (den * path.distance));
    } else {
      double fraction = (time2 - time1) * o->sample_rate;
      potential = v_mul(bperp, sign * constant * fraction / path.distance);
    }
    add_sample(o, w, d2, potential, 1, &counters->zhs_contributions, counters);
    return;
  }
  int nb = (int)(end_bin - start_bin);
  double fraction = fabs(start_bin + 0.5 - (d1 - o->start_time) * o->sample_rate);
  vec3 potential = v_mul(bperp, sign * constant * fraction / (den * path.distance));
  add_sample(o, w, d1, potential, 1, &counters->zhs_contributions, counters);
  for (int index = 1; index < nb; ++index) {
    potential = v_mul(bperp, (subdivided ? 1. : sign) * constant / (den * path.distance));
    add_sample(o, w, d1 + (double)index / o->sample_rate, potential, 1,
               &counters->zhs_contributions, counters);
  }
  fraction = fabs((d2 - o->start_time) * o->sample_rate + 0.5 - end_bin);
  potential = v_mul(bperp, sign * constant * fraction / (den * path.distance));
  add_sample(o, w, d2, potential, 1, &counters->zhs_contributions, counters);
}

/* accumulate_zhs, RadioProjection.inl:482-529 (Fraunhofer subdivision). */
static void zhs_pair(const track_t* t, const c8mi_radio_propagation* prop, uint32_t refinement_u,
                     const obs_t* o, wave_t* w, or_radio_counters* counters) {
  vec3 dest = {o->position[0], o->position[1], o->position[2]};
  vec3 midpoint = v_mul(v_add(t->start, t->end), 0.5);
  path_t mp = propagate(prop, midpoint, dest);
  double u_k = v_dot(t->beta, mp.emit) / t->beta_module;
  double sin2 = 1. - u_k * u_k;
  double wavelength = kC / o->sample_rate;
  double fraunhofer = sin2 * t->length * t->length / mp.distance / wavelength * 2. * kPi;
  double divisor = 1.;
  int nsub = 1;
  if (fraunhofer > 1.) {
    divisor = sqrt(fraunhofer) + 1.;
    nsub = (int)ceil(divisor);
  }
  counters->zhs_subtracks += (uint64_t)nsub;
  vec3 spatial_step = v_div(t->displacement, divisor);
  double time_step = t->duration / divisor;
  int refinement = (int)refinement_u;
  vec3 rs = v_div(spatial_step, (double)refinement);
  double rt = time_step / (double)refinement;
  vec3 p1 = t->start;
  double time1 = t->t0;
  for (int i = 0; i < nsub; ++i) {
    for (int r = 0; r < refinement; ++r) {
      vec3 p2 = v_add(p1, rs);
      double time2 = time1 + rt;
      zhs_segment(p1, p2, time1, time2, t->beta, t->constant, prop, o, w, counters,
                  nsub > 1 || refinement > 1);
      p1 = p2;
      time1 = time2;
    }
  }
  if (refinement > 1) counters->zhs_subtracks += (uint64_t)(nsub * (refinement - 1));
}

void oracle_radio_build_flat_table(const double* refractivity, uint64_t n,
                                   double inverse_step, double* integrated,
                                   double slopes[4]) {
  integrated[0] = refractivity[0] * inverse_step;
  for (uint64_t i = 1; i < n; ++i) integrated[i] = integrated[i - 1] + refractivity[i] * inverse_step;
  uint64_t last = n - 1;
  slopes[0] = (refractivity[10] - refractivity[0]) / 10.;
  slopes[1] = (integrated[10] - integrated[0]) / 10.;
  slopes[2] = (refractivity[last] - refractivity[last - 10]) / 10.;
  slopes[3] = (integrated[last] - integrated[last - 10]) / 10.;
}

static void make_observers(const c8mi_radio_observer* in, uint64_t count, double limit,
                           int zhs_potential, obs_t* out, uint64_t* total_bins) {
  uint64_t total = 0;
  for (uint64_t i = 0; i < count; ++i) {
    for (int a = 0; a < 3; ++a) out[i].position[a] = in[i].position_m[a];
    out[i].start_time = in[i].start_time_s;
    out[i].duration = in[i].duration_s;
    out[i].sample_rate = in[i].sample_rate_hz;
    double potential_factor = zhs_potential ? in[i].sample_rate_hz : 1.;
    out[i].fixed_scale = kHeadroom * potential_factor / limit;
    out[i].nbins = in[i].number_of_bins;
    out[i].offset = total;
    total += in[i].number_of_bins;
  }
  *total_bins = total;
}

static void merge_fixed(int64_t* dst, const int64_t* src, uint64_t n, uint64_t* overflows) {
  for (uint64_t i = 0; i < n; ++i) {
    int64_t previous = dst[i], increment = src[i];
    dst[i] = (int64_t)((uint64_t)previous + (uint64_t)increment);
    if ((increment > 0 && previous > INT64_MAX - increment) ||
        (increment < 0 && previous < INT64_MIN - increment))
      *overflows += 1;
  }
}

int oracle_radio_project(const c8mi_radio_propagation* propagation,
                         uint32_t zhs_refinement, double fixed_point_limit,
                         const c8mi_radio_observer* observers, uint64_t observer_count,
                         int coreas, int zhs, const double* tracks, uint64_t stride,
                         uint64_t track_count, int64_t* coreas_fixed,
                         double* coreas_float, int64_t* zhs_fixed, double* zhs_float,
                         or_radio_counters* counters, int threads) {
  if (zhs_refinement < 1 || zhs_refinement > 64) return 1;
  obs_t* co = (obs_t*)malloc(sizeof(obs_t) * (observer_count ? observer_count : 1));
  obs_t* zo = (obs_t*)malloc(sizeof(obs_t) * (observer_count ? observer_count : 1));
  uint64_t total = 0;
  make_observers(observers, observer_count, fixed_point_limit, 0, co, &total);
  make_observers(observers, observer_count, fixed_point_limit, 1, zo, &total);
  memset(counters, 0, sizeof(*counters));
  int nthreads = threads < 1 ? 1 : threads;
#ifndef _OPENMP
  nthreads = 1;
#endif
  if (nthreads == 1) {
    wave_t cw = {coreas_fixed, coreas_float, total};
    wave_t zw = {zhs_fixed, zhs_float, total};
    for (uint64_t i = 0; i < track_count; ++i) {
      track_t t;
      if (!make_track(tracks, stride, i, &t)) continue;
      counters->valid_tracks += 1;
      for (uint64_t k = 0; k < observer_count; ++k) {
        if (coreas) coreas_pair(&t, propagation, &co[k], &cw, counters);
        if (zhs) zhs_pair(&t, propagation, zhs_refinement, &zo[k], &zw, counters);
      }
    }
  } else {
#ifdef _OPENMP
    uint64_t n3 = 3 * total;
    or_radio_counters* tc = (or_radio_counters*)calloc((size_t)nthreads, sizeof(or_radio_counters));
#pragma omp parallel num_threads(nthreads)
    {
      int tid = omp_get_thread_num();
      wave_t cw = {coreas_fixed ? (int64_t*)calloc(n3, 8) : NULL,
                   coreas_float ? (double*)calloc(n3, 8) : NULL, total};
      wave_t zw = {zhs_fixed ? (int64_t*)calloc(n3, 8) : NULL,
                   zhs_float ? (double*)calloc(n3, 8) : NULL, total};
      or_radio_counters* c = &tc[tid];
#pragma omp for schedule(dynamic, 256)
      for (uint64_t i = 0; i < track_count; ++i) {
        track_t t;
        if (!make_track(tracks, stride, i, &t)) continue;
        c->valid_tracks += 1;
        for (uint64_t k = 0; k < observer_count; ++k) {
          if (coreas) coreas_pair(&t, propagation, &co[k], &cw, c);
          if (zhs) zhs_pair(&t, propagation, zhs_refinement, &zo[k], &zw, c);
        }
      }
#pragma omp critical
      {
        if (cw.fixed) merge_fixed(coreas_fixed, cw.fixed, n3, &counters->fixed_point_overflows);
        if (zw.fixed) merge_fixed(zhs_fixed, zw.fixed, n3, &counters->fixed_point_overflows);
        if (cw.floating)
          for (uint64_t j = 0; j < n3; ++j) coreas_float[j] += cw.floating[j];
        if (zw.floating)
          for (uint64_t j = 0; j < n3; ++j) zhs_float[j] += zw.floating[j];
      }
      free(cw.fixed); free(cw.floating); free(zw.fixed); free(zw.floating);
    }
    for (int t = 0; t < nthreads; ++t) {
      counters->coreas_contributions += tc[t].coreas_contributions;
      counters->zhs_contributions += tc[t].zhs_contributions;
      counters->zhs_subtracks += tc[t].zhs_subtracks;
      counters->valid_tracks += tc[t].valid_tracks;
      counters->fixed_point_overflows += tc[t].fixed_point_overflows;
    }
    free(tc);
#endif
  }
  free(co);
  free(zo);
  return 0;
}

double oracle_radio_nearest_bin(double scaled_time) { return nearest_bin(scaled_time); }

double oracle_radio_doppler(double index, const double beta[3], const double emit[3],
                            int device_variant) {
  vec3 b = {beta[0], beta[1], beta[2]}, e = {emit[0], emit[1], emit[2]};
  return doppler(index, b, e, device_variant);
}

void oracle_radio_propagate(const c8mi_radio_propagation* propagation,
                            const double source[3], const double destination[3],
                            double out[7]) {
  vec3 s = {source[0], source[1], source[2]};
  vec3 d = {destination[0], destination[1], destination[2]};
  path_t p = propagate(propagation, s, d);
  out[0] = p.time;
  out[1] = p.n_source;
  out[2] = p.n_destination;
  out[3] = p.distance;
  out[4] = p.emit.x;
  out[5] = p.emit.y;
  out[6] = p.emit.z;
}

void oracle_radio_separate_same_bin(double* start_time, double* end_time,
                                    double grid_resolution, int signed_order) {
  separate_same_bin(start_time, end_time, grid_resolution, signed_order);
}

void oracle_radio_decode(const int64_t* fixed, uint64_t total_bins,
                         const c8mi_radio_observer* observers, uint64_t observer_count,
                         double fixed_point_limit, int zhs_potential, double* out) {
  uint64_t offset = 0;
  for (uint64_t k = 0; k < observer_count; ++k) {
    double potential_factor = zhs_potential ? observers[k].sample_rate_hz : 1.;
    double inverse = fixed_point_limit / (kHeadroom * potential_factor);
    for (uint64_t b = 0; b < observers[k].number_of_bins; ++b) {
      uint64_t j = offset + b;
      out[j] = (double)fixed[j] * inverse;
      out[total_bins + j] = (double)fixed[total_bins + j] * inverse;
      out[2 * total_bins + j] = (double)fixed[2 * total_bins + j] * inverse;
    }
    offset += observers[k].number_of_bins;
  }
}
