// CoREAS / ZHS endpoint projection for gfx950.
//
// Physics: corsika/detail/kokkos/radio/algorithms/RadioProjection.inl (propagate
// :170-238, accumulate_co_reas :338-409, accumulate_zhs_segment :411-480,
// accumulate_zhs :482-529, add_sample :262-301). Every floating-point expression
// keeps the reference's operation order and the file is built without
// contraction, so the int64 fixed-point waveforms equal the FP64 restatement's.
//
// MI355X design (differs from the reference's track-tile x observer-tile teams,
// RadioAccumulator.inl:15-85):
//  * one workgroup = one observer x one chunk of tracks x one window tile; the
//    observer's whole waveform tile lives in LDS as int64 fixed point, so the
//    per-sample atomics are ds_add_u64, and HBM sees one flush per workgroup;
//  * ZHS middle bins all receive the same increment, so they are added to an
//    LDS difference array (two atomics per component instead of one per bin)
//    and recovered by a block prefix scan before the flush; integer addition
//    makes this bit-identical to the per-bin loop;
//  * threads stride over tracks (coalesced 9 x FP64 structure-of-arrays loads);
//  * blockIdx -> (chunk, observer) is XCD-aware: the eight XCDs take disjoint
//    chunk sets and walk all observers of a chunk back to back, so a chunk is
//    fetched from HBM once per XCD and re-read from that XCD's L2.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <type_traits>

#include "radio_device.hpp"
#include "radio_launch.hpp"
#include "../runtime.hpp"

namespace c8mi::radio {
namespace {

constexpr double kC = 299792458.;
constexpr double kE = 1.6021766208e-19;
constexpr double kEps0 = 8.8541878128e-12;
constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kEmConstant = 1. / (4. * kPi * kEps0 * kC);
constexpr double kChargeEm = kE * kEmConstant;
constexpr double kCoreasThreshold = 1.e-3;
constexpr double kIntLimit = 9223372036854775808.;  // 2^63

constexpr int kXcds = 8;

struct V3 {
  double x, y, z;
};
__device__ __forceinline__ V3 vadd(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 vmul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 vdiv(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 vcross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

__device__ __forceinline__ double nearest_bin(double scaled) {
  double const integral = floor(scaled);
  return integral + (scaled - integral >= 0.5 ? 1. : 0.);
}

// Device form of co_reas_doppler: error-free products and sums when the
// ordinary expression cancels to exactly zero.
__device__ __forceinline__ double doppler(double index, V3 beta, V3 emit) {
  double const ordinary = 1. - index * vdot(beta, emit);
  if (ordinary != 0.) return ordinary;
  double const px = beta.x * emit.x, py = beta.y * emit.y, pz = beta.z * emit.z;
  double const ex = fma(beta.x, emit.x, -px);
  double const ey = fma(beta.y, emit.y, -py);
  double const ez = fma(beta.z, emit.z, -pz);
  double const xy = px + py;
  double const xyv = xy - px;
  double const xye = (px - (xy - xyv)) + (py - xyv);
  double const xyz = xy + pz;
  double const xyzv = xyz - xy;
  double const xyze = (xy - (xyz - xyzv)) + (pz - xyzv);
  double const low = ex + ey + ez + xye + xyze;
  double const product = index * xyz;
  double const perr = fma(index, xyz, -product) + index * low;
  return (1. - product) - perr;
}

// ---- arithmetic of the two precision modes --------------------------------
//
// Exact: the reference's FP64 expressions in its operation order (correctly
// rounded IEEE division and square root, no contraction): the fixed-point
// waveforms equal the FP64 restatement's bit for bit.
// Fast: FP64 throughout, but every division and square root of the pair loop
// becomes a hardware reciprocal / reciprocal square root refined by two Newton
// steps (about one ulp), and the inverse distance is shared by the emission
// direction and the amplitudes. Arrival times must stay at FP64 accuracy: an
// endpoint sample is a spike in one bin, and a delay error of even 1e-15 s at
// 15 km (an FP32 refractivity ratio) moves a few samples per million across a
// bin edge, which shows up as per-bin errors of order 1e-3.

__device__ __forceinline__ double recip_fast(double b) {
  double const r0 = __builtin_amdgcn_rcp(b);
  double e = fma(-b, r0, 1.);
  double r = fma(r0, e, r0);
  e = fma(-b, r, 1.);
  r = fma(r, e, r);
  // 0 and +-inf: keep the hardware result (+-inf, +-0) instead of NaN.
  return isfinite(r0) && r0 != 0. ? r : r0;
}

// Returns sqrt(x) and 1/sqrt(x) for x > 0 (NaN propagates).
__device__ __forceinline__ double rsqrt_fast(double x, double& root) {
  double y = __builtin_amdgcn_rsq(x);
  double h = x * y;
  double e = fma(-h, y, 1.);
  y = fma(0.5 * y, e, y);
  h = x * y;
  e = fma(-h, y, 1.);
  y = fma(0.5 * y, e, y);
  root = x * y;
  return y;
}

struct Track {
  V3 start, end, disp, beta;
  double t0, t1, duration, length, beta_module, constant;
  double inv_beta_module;  // fast mode only
};

template <bool Exact>
__device__ __forceinline__ bool load_track(double const* __restrict__ tracks,
                                           std::uint64_t stride, std::uint64_t i, Track& t) {
  t.start = {tracks[i], tracks[stride + i], tracks[2 * stride + i]};
  t.end = {tracks[3 * stride + i], tracks[4 * stride + i], tracks[5 * stride + i]};
  t.t0 = tracks[6 * stride + i];
  t.t1 = tracks[7 * stride + i];
  double const signed_weight = tracks[8 * stride + i];
  t.duration = t.t1 - t.t0;
  t.disp = vsub(t.end, t.start);
  if constexpr (Exact) {
    t.length = sqrt(vdot(t.disp, t.disp));
    if (!(t.duration > 0.) || !(t.length > 0.)) return false;
    t.beta = vdiv(t.disp, kC * t.duration);
    t.beta_module = t.length / (kC * t.duration);
  } else {
    double const l2 = fma(t.disp.x, t.disp.x, fma(t.disp.y, t.disp.y, t.disp.z * t.disp.z));
    if (!(t.duration > 0.) || !(l2 > 0.)) return false;
    double const inv_length = rsqrt_fast(l2, t.length);
    double const c_duration = kC * t.duration;
    double const inv_c_duration = recip_fast(c_duration);
    t.beta = vmul(t.disp, inv_c_duration);
    t.beta_module = t.length * inv_c_duration;
    t.inv_beta_module = c_duration * inv_length;
  }
  t.constant = kChargeEm * signed_weight;
  return true;
}

// Observer quantities that do not depend on the track, evaluated once per
// thread with the same expressions propagate() and add_sample() use.
struct Obs {
  V3 position;
  double start, duration, rate, grid, wavelength, two_pi_over_wavelength;
  double nbins_d;
  std::int32_t nbins;
  std::uint64_t offset;
  double destination_height;
  std::uint64_t destination_index;
  double n_destination_table, integrated_destination;
};

struct Path {
  double time, n_source, n_destination, distance;
  double inv_distance;  // fast mode only
  V3 emit;
};

template <bool Exact>
__device__ __forceinline__ Path propagate(PropagationDev const& p, V3 source, Obs const& o) {
  Path r;
  V3 const disp = vsub(o.position, source);
  if constexpr (Exact) {
    r.distance = sqrt(vdot(disp, disp));
    r.emit = vdiv(disp, r.distance);
  } else {
    double const r2 = fma(disp.x, disp.x, fma(disp.y, disp.y, disp.z * disp.z));
    r.inv_distance = rsqrt_fast(r2, r.distance);
    r.emit = vmul(disp, r.inv_distance);
  }
  double const distance = r.distance;
  if (p.homogeneous_index > 0.) {
    double const index = p.homogeneous_index;
    if constexpr (Exact) r.time = index * distance / kC;
    else r.time = index * distance * (1. / kC);
    r.n_source = index;
    r.n_destination = index;
    return r;
  }
  double const source_height = (source.z - p.minimum_height) * p.inverse_step;
  double const destination_height = o.destination_height;
  std::uint64_t const last = p.table_size - 1;
  double n_source = 1.;
  double integrated_source = 1.;
  double n_destination = o.n_destination_table;
  double height = 1.;
  if (source_height + 0.5 >= static_cast<double>(last)) {
    n_source = p.refractivity[last] +
               p.slope_n_upper * fabs((source.z - p.maximum_height) * p.inverse_step) + 1.;
    integrated_source =
        p.integrated[last] + p.slope_i_upper * fabs(p.maximum_height - source.z);
    height = source_height - destination_height;
  } else if (source_height + 0.5 < static_cast<double>(last) && source_height > 0.) {
    // The table has fewer than 2^32 entries (checked on the host).
    std::uint32_t const source_index = static_cast<std::uint32_t>(source_height + 0.5);
    n_source = p.refractivity[source_index] + 1.;
    integrated_source = p.integrated[source_index];
    height = static_cast<double>(static_cast<std::int64_t>(source_index) -
                                 static_cast<std::int64_t>(o.destination_index)) *
             p.step;
    if (height == 0.) height = 1.;
  } else if (source_height == 0.) {
    n_source = 1.;
    n_destination = 1.;
    integrated_source = p.integrated[0];
    height = destination_height - source_height;
  } else {
    n_source = p.refractivity[0] + p.slope_n_lower * fabs(source_height) + 1.;
    integrated_source = p.integrated[0] + p.slope_i_lower * fabs(source_height);
    n_destination = 1.;
    height = destination_height - fabs(source_height);
  }
  if constexpr (Exact) {
    r.time = (1. + (integrated_source - o.integrated_destination) / height) * distance / kC;
  } else {
    double const ratio = (integrated_source - o.integrated_destination) * recip_fast(height);
    r.time = (1. + ratio) * (distance * (1. / kC));
  }
  r.n_source = n_source;
  r.n_destination = n_destination;
  return r;
}

// Per-thread tallies, folded into the block and then the global counters.
struct Tally {
  unsigned long long coreas = 0, zhs = 0, subtracks = 0, valid = 0, overflow = 0;
  double abs_sum = 0.;  // sum of |increments|: bounds every LDS partial sum
};

// LDS accumulator for one mode: `Fixed` selects int64 fixed point (the
// production default) or FP64. The ZHS tile stores a difference array.
template <bool Fixed>
struct Tile {
  using Acc = std::conditional_t<Fixed, long long, double>;
  Acc* data;             // 3 components x stride
  std::uint32_t stride;  // bins per component row
  std::int32_t begin;    // first global bin of this tile
  std::int32_t bins;     // bins owned by this tile
};

// add_radio_fixed_point on an LDS cell; returns false on a non-representable
// contribution (the reference's overflow case), in which case nothing is added.
// Below 2^51 the round-to-nearest-even integer comes from adding 1.5 * 2^52
// and reading the low mantissa bits: the same value as (long long) rint().
__device__ __forceinline__ bool fixed_increment(double value, double scale, long long& inc,
                                                double& abs_sum) {
  double const scaled = value * scale;
  double const magnitude = fabs(scaled);
  if (magnitude < 2251799813685248.) {  // 2^51
    double const shifted = scaled + 6755399441055744.;  // 1.5 * 2^52
    inc = __double_as_longlong(shifted) - 0x4338000000000000LL;
  } else {
    if (!isfinite(scaled) || magnitude >= kIntLimit) return false;
    inc = static_cast<long long>(rint(scaled));
  }
  abs_sum += magnitude;
  return true;
}

template <class T>
__device__ __forceinline__ void lds_add(T* address, T value) {
  if constexpr (std::is_same_v<T, long long>)
    atomicAdd(reinterpret_cast<unsigned long long*>(address),
              static_cast<unsigned long long>(value));
  else
    atomicAdd(address, value);
}

// Bin of `time` per observer_nearest_bin; -1 when outside [0, nbins).
__device__ __forceinline__ std::int32_t sample_bin(Obs const& o, double time) {
  double const bin_value = nearest_bin((time - o.start) * o.rate);
  if (!(bin_value >= 0. && bin_value < o.nbins_d)) return -1;
  return static_cast<std::int32_t>(bin_value);
}

// CoREAS sample (add_sample with time gating, IsVectorPotential = false).
template <bool Fixed>
__device__ __forceinline__ void add_field_sample(Obs const& o, double scale,
                                                 Tile<Fixed> const& tile, double time, V3 c,
                                                 Tally& tally) {
  if (time < o.start || time > o.start + o.duration) return;
  std::int32_t const bin = sample_bin(o, time);
  if (bin < 0) return;
  std::int32_t const local = bin - tile.begin;
  if (local < 0 || local >= tile.bins) return;
  if constexpr (Fixed) {
    long long ix = 0, iy = 0, iz = 0;
    bool const vx = fixed_increment(c.x, scale, ix, tally.abs_sum);
    bool const vy = fixed_increment(c.y, scale, iy, tally.abs_sum);
    bool const vz = fixed_increment(c.z, scale, iz, tally.abs_sum);
    if (vx) lds_add(tile.data + local, ix);
    if (vy) lds_add(tile.data + tile.stride + local, iy);
    if (vz) lds_add(tile.data + 2 * tile.stride + local, iz);
    // Counter updates are value selects: a branch between two tally fields
    // can become a store through a selected pointer, which pins the tally to
    // scratch memory.
    bool const valid = vx && vy && vz;
    tally.overflow += valid ? 0ull : 1ull;
    tally.coreas += valid ? 1ull : 0ull;
    return;
  } else {
    lds_add(tile.data + local, c.x);
    lds_add(tile.data + tile.stride + local, c.y);
    lds_add(tile.data + 2 * tile.stride + local, c.z);
  }
  ++tally.coreas;
}

// Adds `c` to the ZHS bins [first, last] (inside [0, nbins)) through the
// difference array, clipped to the tile. Returns the number of this tile's
// bins that received it.
template <bool Fixed>
__device__ __forceinline__ std::int32_t add_potential_range(double scale,
                                                            Tile<Fixed> const& tile,
                                                            std::int32_t first,
                                                            std::int32_t last, V3 c,
                                                            Tally& tally, bool& valid) {
  std::int32_t const lo = (first > tile.begin ? first : tile.begin) - tile.begin;
  std::int32_t const hi =
      (last + 1 < tile.begin + tile.bins ? last + 1 : tile.begin + tile.bins) - tile.begin;
  valid = true;
  if (lo >= hi) return 0;
  if constexpr (Fixed) {
    long long ix = 0, iy = 0, iz = 0;
    bool const vx = fixed_increment(c.x, scale, ix, tally.abs_sum);
    bool const vy = fixed_increment(c.y, scale, iy, tally.abs_sum);
    bool const vz = fixed_increment(c.z, scale, iz, tally.abs_sum);
    if (vx) {
      lds_add(tile.data + lo, ix);
      lds_add(tile.data + hi, -ix);
    }
    if (vy) {
      lds_add(tile.data + tile.stride + lo, iy);
      lds_add(tile.data + tile.stride + hi, -iy);
    }
    if (vz) {
      lds_add(tile.data + 2 * tile.stride + lo, iz);
      lds_add(tile.data + 2 * tile.stride + hi, -iz);
    }
    valid = vx && vy && vz;
  } else {
    lds_add(tile.data + lo, c.x);
    lds_add(tile.data + hi, -c.x);
    lds_add(tile.data + tile.stride + lo, c.y);
    lds_add(tile.data + tile.stride + hi, -c.y);
    lds_add(tile.data + 2 * tile.stride + lo, c.z);
    lds_add(tile.data + 2 * tile.stride + hi, -c.z);
  }
  return hi - lo;
}

// ZHS sample (add_sample with IsVectorPotential = true): no time gate.
template <bool Fixed>
__device__ __forceinline__ void add_potential_sample(Obs const& o, double scale,
                                                     Tile<Fixed> const& tile, double time, V3 c,
                                                     Tally& tally) {
  std::int32_t const bin = sample_bin(o, time);
  if (bin < 0) return;
  bool valid = true;
  std::int32_t const added = add_potential_range(scale, tile, bin, bin, c, tally, valid);
  unsigned long long const n = static_cast<unsigned long long>(added);
  tally.zhs += valid ? n : 0ull;
  tally.overflow += valid ? 0ull : n;
}

// separate_same_bin_endpoints (RadioProjection.inl:303-336); `to_bins` is the
// reference's division by the grid (exact) or the product with the rate (fast).
template <bool Exact>
__device__ __forceinline__ double to_bins(double time, Obs const& o) {
  if constexpr (Exact) return time / o.grid;
  else return time * o.rate;
}

template <bool Exact>
__device__ __forceinline__ void separate_same_bin(double& start_time, double& end_time,
                                                  Obs const& o, bool signed_order) {
  double const s = to_bins<Exact>(start_time, o);
  double const e = to_bins<Exact>(end_time, o);
  if (static_cast<long long>(nearest_bin(s)) != static_cast<long long>(nearest_bin(e))) return;
  double const grid = o.grid;
  double const sf = s - floor(s);
  double const ef = e - floor(e);
  bool const forward = !signed_order || end_time >= start_time;
  if (forward) {
    if (sf >= 0.5 && ef >= 0.5) start_time -= grid;
    else if (sf < 0.5 && ef < 0.5) end_time += grid;
    else if (ef >= 1. - sf) end_time += grid;
    else start_time -= grid;
  } else {
    if (sf >= 0.5 && ef >= 0.5) end_time -= grid;
    else if (sf < 0.5 && ef < 0.5) start_time += grid;
    else if (sf >= 1. - ef) start_time += grid;
    else end_time -= grid;
  }
}

// t.constant * rate / (doppler * R)
template <bool Exact>
__device__ __forceinline__ double field_scale(Track const& t, Obs const& o, double doppler_factor,
                                              Path const& path) {
  if constexpr (Exact) return t.constant * o.rate / (doppler_factor * path.distance);
  else return (t.constant * o.rate) * path.inv_distance * recip_fast(doppler_factor);
}

template <bool Exact, bool Fixed>
__device__ __forceinline__ void coreas_pair(Track const& t, PropagationDev const& p, Obs const& o,
                            double scale, Tile<Fixed> const& tile, Tally& tally) {
  Path const ps = propagate<Exact>(p, t.start, o);
  Path const pe = propagate<Exact>(p, t.end, o);
  double const pre = doppler(ps.n_source, t.beta, ps.emit);
  double const post = doppler(pe.n_source, t.beta, pe.emit);
  double start_receive = t.t0 + ps.time;
  double end_receive = t.t1 + pe.time;
  if (ps.n_destination > 1. &&
      (fabs(pre) < kCoreasThreshold || fabs(post) < kCoreasThreshold)) {
    V3 const midpoint = vmul(vadd(t.start, t.end), 0.5);
    double const middle_time = (t.t0 + t.t1) * 0.5;
    Path const pm = propagate<Exact>(p, midpoint, o);
    double const middle_receive = middle_time + pm.time;
    double const md = doppler(pm.n_source, t.beta, pm.emit);
    V3 field_start =
        vmul(vcross(pm.emit, vcross(pm.emit, t.beta)), field_scale<Exact>(t, o, md, pm));
    V3 field_end = vmul(field_start, -1.);
    double delta;
    if constexpr (Exact) delta = t.length / (kC * t.beta_module) * fabs(md);
    else delta = t.duration * fabs(md);  // length / (c beta) is the duration
    if (start_receive < end_receive) {
      start_receive = middle_receive - 0.5 * delta;
      end_receive = middle_receive + 0.5 * delta;
    } else {
      start_receive = middle_receive + 0.5 * delta;
      end_receive = middle_receive - 0.5 * delta;
    }
    delta = end_receive - start_receive;
    if (fabs(delta) < o.grid) {
      double const shrink = fabs(to_bins<Exact>(delta, o));
      field_start = vmul(field_start, shrink);
      field_end = vmul(field_end, shrink);
      separate_same_bin<Exact>(start_receive, end_receive, o, true);
    }
    add_field_sample(o, scale, tile, start_receive, field_start, tally);
    add_field_sample(o, scale, tile, end_receive, field_end, tally);
    return;
  }
  V3 field_start =
      vmul(vcross(ps.emit, vcross(ps.emit, t.beta)), field_scale<Exact>(t, o, pre, ps));
  V3 field_end =
      vmul(vcross(pe.emit, vcross(pe.emit, t.beta)), -field_scale<Exact>(t, o, post, pe));
  if (pre < 1.e-9 || post < 1.e-9) {
    double const delta = end_receive - start_receive;
    if (fabs(delta) < o.grid) {
      double const shrink = fabs(to_bins<Exact>(delta, o));
      field_start = vmul(field_start, shrink);
      field_end = vmul(field_end, shrink);
      separate_same_bin<Exact>(start_receive, end_receive, o, false);
    }
  }
  add_field_sample(o, scale, tile, start_receive, field_start, tally);
  add_field_sample(o, scale, tile, end_receive, field_end, tally);
}

template <bool Exact, bool Fixed>
__device__ __forceinline__ void zhs_segment(V3 p1, V3 p2, double time1, double time2, Track const& t,
                            PropagationDev const& p, Obs const& o, double scale,
                            Tile<Fixed> const& tile, Tally& tally, bool subdivided,
                            bool use_cached, Path cached_midpoint_path) {
  // The cached path travels by value: a pointer to the caller's local would
  // pin it to scratch memory.
  Path const path = use_cached ? cached_midpoint_path
                               : propagate<Exact>(p, vmul(vadd(p1, p2), 0.5), o);
  double const n_source = path.n_source;
  double const bk = vdot(t.beta, path.emit);
  double const middle_time = (time1 + time2) * 0.5;
  double d1 = time1 + path.time - n_source * bk * (time1 - middle_time);
  double d2 = time2 + path.time - n_source * bk * (time2 - middle_time);
  double sign = 1.;
  if (d1 > d2) {
    d1 = time2 + path.time - n_source * bk * (time2 - middle_time);
    d2 = time1 + path.time - n_source * bk * (time1 - middle_time);
    sign = -1.;
  }
  double const x1 = (d1 - o.start) * o.rate;
  double const start_bin = floor(x1 + 0.5);
  double const end_bin = floor((d2 - o.start) * o.rate + 0.5);
  V3 const bperp = vcross(path.emit, vcross(t.beta, path.emit));
  double const den = 1. - n_source * bk;
  // constant / (den * R)
  double k;
  if constexpr (Exact) k = 0.;  // divisions below keep the reference's form
  else k = t.constant * path.inv_distance * recip_fast(den);
  if (start_bin == end_bin) {
    V3 potential;
    if (fabs(den) > 1.e-15) {
      double const fraction = fabs(d2 * o.rate - d1 * o.rate);
      if constexpr (Exact)
        potential = vmul(bperp, sign * t.constant * fraction / (den * path.distance));
      else
        potential = vmul(bperp, sign * k * fraction);
    } else {
      double const fraction = (time2 - time1) * o.rate;
      if constexpr (Exact)
        potential = vmul(bperp, sign * t.constant * fraction / path.distance);
      else
        potential = vmul(bperp, sign * t.constant * fraction * path.inv_distance);
    }
    add_potential_sample(o, scale, tile, d2, potential, tally);
    return;
  }
  int const nb = static_cast<int>(end_bin - start_bin);
  double fraction = fabs(start_bin + 0.5 - x1);
  V3 potential;
  if constexpr (Exact)
    potential = vmul(bperp, sign * t.constant * fraction / (den * path.distance));
  else
    potential = vmul(bperp, sign * k * fraction);
  add_potential_sample(o, scale, tile, d1, potential, tally);
  if (nb > 1) {
    V3 middle;
    if constexpr (Exact)
      middle = vmul(bperp, (subdivided ? 1. : sign) * t.constant / (den * path.distance));
    else
      middle = vmul(bperp, (subdivided ? 1. : sign) * k);
    // Bins of the first and last middle sample, exactly as the per-bin loop
    // computes them; in between the bins are consecutive.
    double b_first, b_last;
    if constexpr (Exact) {
      b_first = nearest_bin((d1 + 1. / o.rate - o.start) * o.rate);
      b_last = nearest_bin((d1 + static_cast<double>(nb - 1) / o.rate - o.start) * o.rate);
    } else {
      b_first = nearest_bin(x1 + 1.);
      b_last = nearest_bin(x1 + static_cast<double>(nb - 1));
    }
    if (b_last - b_first == static_cast<double>(nb - 2)) {
      double const lo = b_first < 0. ? 0. : b_first;
      double const hi = b_last < o.nbins_d - 1. ? b_last : o.nbins_d - 1.;
      if (lo <= hi) {
        bool valid = true;
        std::int32_t const added =
            add_potential_range(scale, tile, static_cast<std::int32_t>(lo),
                                static_cast<std::int32_t>(hi), middle, tally, valid);
        unsigned long long const n = static_cast<unsigned long long>(added);
        tally.zhs += valid ? n : 0ull;
        tally.overflow += valid ? 0ull : n;
      }
    } else {
      for (int index = 1; index < nb; ++index)
        add_potential_sample(o, scale, tile, d1 + static_cast<double>(index) / o.rate, middle,
                             tally);
    }
  }
  fraction = fabs((d2 - o.start) * o.rate + 0.5 - end_bin);
  if constexpr (Exact)
    potential = vmul(bperp, sign * t.constant * fraction / (den * path.distance));
  else
    potential = vmul(bperp, sign * k * fraction);
  add_potential_sample(o, scale, tile, d2, potential, tally);
}

template <bool Exact, bool Fixed>
__device__ __forceinline__ void zhs_pair(Track const& t, PropagationDev const& p, std::uint32_t refinement_u,
                         Obs const& o, double scale, Tile<Fixed> const& tile, Tally& tally,
                         bool count_subtracks) {
  V3 const midpoint = vmul(vadd(t.start, t.end), 0.5);
  Path const mp = propagate<Exact>(p, midpoint, o);
  double u_k, fraunhofer;
  if constexpr (Exact) {
    u_k = vdot(t.beta, mp.emit) / t.beta_module;
    double const sin2 = 1. - u_k * u_k;
    fraunhofer = sin2 * t.length * t.length / mp.distance / o.wavelength * 2. * kPi;
  } else {
    u_k = vdot(t.beta, mp.emit) * t.inv_beta_module;
    double const sin2 = 1. - u_k * u_k;
    fraunhofer = sin2 * t.length * t.length * mp.inv_distance * o.two_pi_over_wavelength;
  }
  double divisor = 1.;
  int nsub = 1;
  if (fraunhofer > 1.) {
    divisor = sqrt(fraunhofer) + 1.;
    nsub = static_cast<int>(ceil(divisor));
  }
  int const refinement = static_cast<int>(refinement_u);
  if (count_subtracks) {
    tally.subtracks += static_cast<unsigned long long>(nsub);
    if (refinement > 1)
      tally.subtracks += static_cast<unsigned long long>(nsub * (refinement - 1));
  }
  bool const subdivided = nsub > 1 || refinement > 1;
  // Unsubdivided, divisor == refinement == 1 and x / 1. is x exactly: the
  // single segment runs from start to start + displacement over the full
  // duration, so the divisions are skipped. Its midpoint is the track
  // midpoint whenever start + displacement == end bit for bit, and the
  // propagation above is then reused unchanged.
  V3 rs = t.disp;
  double rt = t.duration;
  bool use_cached = false;
  if (subdivided) {
    V3 const spatial_step = vdiv(t.disp, divisor);
    double const time_step = t.duration / divisor;
    rs = vdiv(spatial_step, static_cast<double>(refinement));
    rt = time_step / static_cast<double>(refinement);
  } else {
    V3 const p2 = vadd(t.start, t.disp);
    use_cached = __double_as_longlong(p2.x) == __double_as_longlong(t.end.x) &&
                 __double_as_longlong(p2.y) == __double_as_longlong(t.end.y) &&
                 __double_as_longlong(p2.z) == __double_as_longlong(t.end.z);
  }
  int const segments = nsub * refinement;
  V3 p1 = t.start;
  double time1 = t.t0;
  for (int i = 0; i < segments; ++i) {
    V3 const p2 = vadd(p1, rs);
    double const time2 = time1 + rt;
    zhs_segment<Exact>(p1, p2, time1, time2, t, p, o, scale, tile, tally, subdivided, use_cached,
                       mp);
    p1 = p2;
    time1 = time2;
  }
}

__device__ __forceinline__ Obs make_obs(ObserverDev const& d, PropagationDev const& p) {
  Obs o;
  o.position = {d.position[0], d.position[1], d.position[2]};
  o.start = d.start_time;
  o.duration = d.duration;
  o.rate = d.sample_rate;
  o.grid = 1. / d.sample_rate;
  o.wavelength = kC / d.sample_rate;
  o.two_pi_over_wavelength = 2. * kPi / o.wavelength;
  o.nbins = static_cast<std::int32_t>(d.number_of_bins);
  o.nbins_d = static_cast<double>(d.number_of_bins);
  o.offset = d.waveform_offset;
  if (p.homogeneous_index > 0.) {
    o.destination_height = 0.;
    o.destination_index = 0;
    o.n_destination_table = 1.;
    o.integrated_destination = 0.;
  } else {
    o.destination_height = (o.position.z - p.minimum_height) * p.inverse_step;
    o.destination_index = static_cast<std::uint64_t>(o.destination_height + 0.5);
    o.n_destination_table = p.refractivity[o.destination_index] + 1.;
    o.integrated_destination = p.integrated[o.destination_index];
  }
  return o;
}

// Inclusive prefix sum over `count` entries of an LDS row, whole block: each
// thread sums a contiguous run, the run totals are scanned within each wave by
// shuffles and across waves by the first wave (wave_totals: one entry per wave).
template <class T>
__device__ __forceinline__ T wave_inclusive_sum(T x, std::uint32_t lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T const y = __shfl_up(x, d, 64);
    if (static_cast<int>(lane) >= d) x += y;
  }
  return x;
}

template <class T>
__device__ void block_inclusive_scan(T* row, std::uint32_t count, T* wave_totals) {
  std::uint32_t const tid = threadIdx.x;
  std::uint32_t const threads = blockDim.x;
  std::uint32_t const lane = tid & 63u, wave = tid >> 6;
  std::uint32_t const waves = (threads + 63u) >> 6;
  std::uint32_t const per = (count + threads - 1) / threads;
  std::uint32_t const begin = tid * per < count ? tid * per : count;
  std::uint32_t const end = begin + per < count ? begin + per : count;
  T running = 0;
  for (std::uint32_t i = begin; i < end; ++i) {
    running += row[i];
    row[i] = running;
  }
  T const inclusive = wave_inclusive_sum(running, lane);
  if (lane == 63u) wave_totals[wave] = inclusive;
  __syncthreads();
  if (wave == 0) {
    T const total = lane < waves ? wave_totals[lane] : T(0);
    T const before = wave_inclusive_sum(total, lane) - total;
    if (lane < waves) wave_totals[lane] = before;
  }
  __syncthreads();
  T const offset = wave_totals[wave] + (inclusive - running);
  for (std::uint32_t i = begin; i < end; ++i) row[i] += offset;
  __syncthreads();
}

template <bool Coreas, bool Zhs, bool Fixed, bool Exact, int Block>
__global__ void __launch_bounds__(Block, 4) radio_project_kernel(ProjectionLaunch L) {
  using Acc = std::conditional_t<Fixed, long long, double>;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __shared__ Acc scan_scratch[Zhs ? Block / 64 : 1];
  __shared__ unsigned long long block_counts[5];
  __shared__ double block_abs[2];

  std::uint32_t observer, bin_tile;
  std::uint64_t chunk_begin, chunk_end;
  CountersDev* counters = L.counters;
  bool first_observer_tile;
  if (L.segments == nullptr) {
    // XCD-aware decode of blockIdx.x into (chunk, observer, window tile).
    std::uint32_t const linear = blockIdx.x;
    std::uint32_t const xcd = linear % kXcds;
    std::uint32_t const sequence = linear / kXcds;
    std::uint32_t const per_chunk = L.observer_count * L.bin_tiles;
    std::uint32_t const chunk = (sequence / per_chunk) * kXcds + xcd;
    if (chunk >= L.chunks) return;
    std::uint32_t const within = sequence % per_chunk;
    observer = within / L.bin_tiles;
    bin_tile = within % L.bin_tiles;
    chunk_begin = static_cast<std::uint64_t>(chunk) * L.chunk_size;
    chunk_end = chunk_begin + L.chunk_size < L.track_count ? chunk_begin + L.chunk_size : L.track_count;
    first_observer_tile = L.count_valid && observer == 0 && bin_tile == 0;
  } else {
    // Pooled showers: the segment whose items hold blockIdx.x (binary search
    // over item_begin), then (chunk, observer of the slot, window tile).
    unsigned long long const item = blockIdx.x;
    if (item >= *L.item_count) return;
    std::uint32_t lo = 0, hi = *L.segment_count;  // segments[lo].item_begin <= item
    while (hi - lo > 1) {
      std::uint32_t const mid = (lo + hi) / 2;
      if (L.segments[mid].item_begin <= item) lo = mid;
      else hi = mid;
    }
    SegmentDev const seg = L.segments[lo];
    std::uint64_t const local = item - seg.item_begin;
    std::uint64_t const per_chunk =
        static_cast<std::uint64_t>(L.slot_observer_counts[seg.slot]) * L.bin_tiles;
    std::uint64_t const chunk = local / per_chunk;
    std::uint32_t const within = static_cast<std::uint32_t>(local % per_chunk);
    std::uint32_t const slot_observer = within / L.bin_tiles;
    observer = seg.slot * L.slot_observers + slot_observer;
    bin_tile = within % L.bin_tiles;
    chunk_begin = seg.begin + chunk * L.chunk_size;
    chunk_end = chunk_begin + L.chunk_size < seg.end ? chunk_begin + L.chunk_size : seg.end;
    counters = L.counters + seg.slot;
    first_observer_tile = slot_observer == 0 && bin_tile == 0;
  }

  std::uint32_t const W = L.tile_bins;
  Acc* const coreas_lds = reinterpret_cast<Acc*>(lds_raw);
  Acc* const zhs_lds = coreas_lds + (Coreas ? 3 * W : 0);
  std::uint32_t const lds_count = (Coreas ? 3 * W : 0) + (Zhs ? 3 * (W + 1) : 0);
  for (std::uint32_t i = threadIdx.x; i < lds_count; i += blockDim.x) coreas_lds[i] = 0;
  if (threadIdx.x < 5) block_counts[threadIdx.x] = 0;
  if (threadIdx.x < 2) block_abs[threadIdx.x] = 0.;
  __syncthreads();

  PropagationDev const prop = L.propagation;
  std::uint64_t const tile_begin = static_cast<std::uint64_t>(bin_tile) * W;
  // A fused launch has identical CoREAS and ZHS observers (position, window,
  // rate, bins and hence offsets); only their fixed-point scales differ.
  ObserverDev const& view = Coreas ? L.coreas_observers[observer] : L.zhs_observers[observer];
  Obs const o = make_obs(view, prop);
  double const coreas_scale = Coreas ? L.coreas_observers[observer].fixed_scale : 0.;
  double const zhs_scale = Zhs ? L.zhs_observers[observer].fixed_scale : 0.;
  std::uint64_t const nbins = static_cast<std::uint64_t>(o.nbins);
  std::uint32_t const tile_bins =
      tile_begin < nbins ? static_cast<std::uint32_t>(nbins - tile_begin < W ? nbins - tile_begin : W)
                         : 0u;
  Tile<Fixed> const ctile{coreas_lds, W, static_cast<std::int32_t>(tile_begin),
                          static_cast<std::int32_t>(tile_bins)};
  Tile<Fixed> const ztile{zhs_lds, W + 1, static_cast<std::int32_t>(tile_begin),
                          static_cast<std::int32_t>(tile_bins)};

  Tally ctally, ztally;
  for (std::uint64_t i = chunk_begin + threadIdx.x; i < chunk_end; i += blockDim.x) {
    Track t;
    if (!load_track<Exact>(L.tracks, L.stride, i, t)) continue;
    if (first_observer_tile) ++ctally.valid;
    if constexpr (Coreas) {
      if (tile_bins) coreas_pair<Exact>(t, prop, o, coreas_scale, ctile, ctally);
    }
    if constexpr (Zhs) {
      zhs_pair<Exact>(t, prop, L.zhs_refinement, o, zhs_scale, ztile, ztally, bin_tile == 0);
    }
  }

  atomicAdd(&block_counts[0], ctally.coreas);
  atomicAdd(&block_counts[1], ztally.zhs);
  atomicAdd(&block_counts[2], ztally.subtracks);
  atomicAdd(&block_counts[3], ctally.valid);
  atomicAdd(&block_counts[4], ctally.overflow + ztally.overflow);
  if constexpr (Fixed) {
    atomicAdd(&block_abs[0], ctally.abs_sum);
    atomicAdd(&block_abs[1], ztally.abs_sum);
  }
  __syncthreads();

  if constexpr (Zhs) {
    for (int c = 0; c < 3; ++c) block_inclusive_scan(zhs_lds + c * (W + 1), W, scan_scratch);
  }

  // Flush the tile into the global waveforms (checked int64 atomics, as
  // checked_radio_atomic_add; or FP64 atomics in floating mode).
  unsigned long long flush_overflows = 0;
  auto flush = [&](Acc const* lds, std::uint32_t row, std::uint32_t bins, void* global,
                   std::uint64_t total_bins) {
    Acc* const g = static_cast<Acc*>(global);
    for (std::uint32_t idx = threadIdx.x; idx < 3 * bins; idx += blockDim.x) {
      std::uint32_t const c = idx / bins;
      std::uint32_t const b = idx % bins;
      Acc const v = lds[c * row + b];
      if (v == 0) continue;
      std::uint64_t const at = c * total_bins + o.offset + tile_begin + b;
      if constexpr (Fixed) {
        long long const previous = static_cast<long long>(
            atomicAdd(reinterpret_cast<unsigned long long*>(g + at),
                      static_cast<unsigned long long>(v)));
        if ((v > 0 && previous > 0x7fffffffffffffffLL - v) ||
            (v < 0 && previous < (-0x7fffffffffffffffLL - 1) - v))
          ++flush_overflows;
      } else {
        atomicAdd(g + at, v);
      }
    }
  };
  if constexpr (Coreas) flush(coreas_lds, W, tile_bins, L.coreas_waves, L.coreas_total_bins);
  if constexpr (Zhs) flush(zhs_lds, W + 1, tile_bins, L.zhs_waves, L.zhs_total_bins);
  if (flush_overflows) atomicAdd(&block_counts[4], flush_overflows);
  __syncthreads();

  if (threadIdx.x == 0) {
    unsigned long long overflow = block_counts[4];
    // Integer sums are exact modulo 2^64; every LDS partial is bounded by the
    // sum of |increments|, so a block whose bound reaches 2^63 may have lost
    // its result and is reported as an overflow (conservatively).
    if constexpr (Fixed) {
      if (block_abs[0] >= kIntLimit || block_abs[1] >= kIntLimit) ++overflow;
    }
    if (block_counts[0]) atomicAdd(&counters->coreas_contributions, block_counts[0]);
    if (block_counts[1]) atomicAdd(&counters->zhs_contributions, block_counts[1]);
    if (block_counts[2]) atomicAdd(&counters->zhs_subtracks, block_counts[2]);
    if (block_counts[3]) atomicAdd(&counters->valid_tracks, block_counts[3]);
    if (overflow) atomicAdd(&counters->fixed_point_overflows, overflow);
  }
}

// LeptonTransportRecord -> compact track (make_radio_track_kinematics input
// selection, RadioProjection.inl:145-168). Non-e+- records become zero-length
// tracks, which the projection rejects exactly as the reference does.
__global__ void __launch_bounds__(256)
    records_to_tracks_kernel(c8mi_lepton_transport_record const* __restrict__ records,
                             std::uint64_t count, double* __restrict__ tracks,
                             std::uint64_t stride) {
  std::uint64_t const i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  c8mi_lepton_transport_record const& r = records[i];
  int const pid = r.start.pid;
  bool const lepton = pid == C8MI_PID_ELECTRON || pid == C8MI_PID_POSITRON;
  double const w = r.start.weight;
  tracks[i] = lepton ? r.start.position_m[0] : 0.;
  tracks[stride + i] = lepton ? r.start.position_m[1] : 0.;
  tracks[2 * stride + i] = lepton ? r.start.position_m[2] : 0.;
  tracks[3 * stride + i] = lepton ? r.end.position_m[0] : 0.;
  tracks[4 * stride + i] = lepton ? r.end.position_m[1] : 0.;
  tracks[5 * stride + i] = lepton ? r.end.position_m[2] : 0.;
  tracks[6 * stride + i] = lepton ? r.start.time_s : 0.;
  tracks[7 * stride + i] = lepton ? r.end.time_s : 0.;
  tracks[8 * stride + i] = pid == C8MI_PID_ELECTRON ? -w : w;
}

template <bool Coreas, bool Zhs, bool Fixed, bool Exact, int Block>
void launch_typed(ProjectionLaunch const& launch, hipStream_t stream, std::size_t lds_bytes,
                  std::uint32_t grid) {
  auto kernel = radio_project_kernel<Coreas, Zhs, Fixed, Exact, Block>;
  static bool attribute_set = false;
  if (!attribute_set) {
    C8MI_HIP(hipFuncSetAttribute(reinterpret_cast<void const*>(kernel),
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(kMaxDynamicLdsBytes)));
    attribute_set = true;
  }
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(Block), lds_bytes, stream, launch);
  C8MI_HIP(hipGetLastError());
}

template <bool Coreas, bool Zhs, bool Fixed, bool Exact>
void launch_block(ProjectionLaunch const& launch, std::uint32_t block, hipStream_t stream,
                  std::size_t lds_bytes, std::uint32_t grid) {
  switch (block) {
    case 256: launch_typed<Coreas, Zhs, Fixed, Exact, 256>(launch, stream, lds_bytes, grid); break;
    case 512: launch_typed<Coreas, Zhs, Fixed, Exact, 512>(launch, stream, lds_bytes, grid); break;
    case 1024: launch_typed<Coreas, Zhs, Fixed, Exact, 1024>(launch, stream, lds_bytes, grid); break;
    default: throw std::invalid_argument("radio workgroup size must be 256, 512 or 1024");
  }
}

}  // namespace

std::size_t projection_lds_bytes(bool coreas, bool zhs, std::uint32_t tile_bins) {
  std::size_t const words = (coreas ? 3u * tile_bins : 0u) + (zhs ? 3u * (tile_bins + 1u) : 0u);
  return words * sizeof(long long);
}

std::uint32_t projection_grid(ProjectionLaunch const& launch) {
  std::uint64_t const per_chunk =
      static_cast<std::uint64_t>(launch.observer_count) * launch.bin_tiles;
  std::uint64_t const chunk_rounds = (launch.chunks + kXcds - 1) / kXcds;
  return static_cast<std::uint32_t>(chunk_rounds * kXcds * per_chunk);
}

template <bool Coreas, bool Zhs>
void launch_modes(ProjectionLaunch const& launch, bool fixed, bool exact, std::uint32_t block,
                  hipStream_t stream, std::size_t lds, std::uint32_t grid) {
  // FP64 (non-deterministic) accumulation is offered in the exact form only.
  if (!fixed) launch_block<Coreas, Zhs, false, true>(launch, block, stream, lds, grid);
  else if (exact) launch_block<Coreas, Zhs, true, true>(launch, block, stream, lds, grid);
  else launch_block<Coreas, Zhs, true, false>(launch, block, stream, lds, grid);
}

void launch_projection(ProjectionLaunch const& launch, bool coreas, bool zhs, bool fixed,
                       bool exact, std::uint32_t block, hipStream_t stream) {
  if (launch.track_count == 0 || launch.observer_count == 0) return;
  std::size_t const lds = projection_lds_bytes(coreas, zhs, launch.tile_bins);
  std::uint32_t const grid = projection_grid(launch);
  if (coreas && zhs) launch_modes<true, true>(launch, fixed, exact, block, stream, lds, grid);
  else if (coreas) launch_modes<true, false>(launch, fixed, exact, block, stream, lds, grid);
  else if (zhs) launch_modes<false, true>(launch, fixed, exact, block, stream, lds, grid);
}

void launch_pooled_projection(ProjectionLaunch const& launch, bool coreas, bool zhs, bool exact,
                              std::uint32_t block, std::uint32_t grid, hipStream_t stream) {
  if (launch.track_count == 0 || grid == 0) return;
  std::size_t const lds = projection_lds_bytes(coreas, zhs, launch.tile_bins);
  if (coreas && zhs) launch_modes<true, true>(launch, true, exact, block, stream, lds, grid);
  else if (coreas) launch_modes<true, false>(launch, true, exact, block, stream, lds, grid);
  else if (zhs) launch_modes<false, true>(launch, true, exact, block, stream, lds, grid);
}

void launch_records_to_tracks(c8mi_lepton_transport_record const* records, std::uint64_t count,
                              double* tracks, std::uint64_t stride, hipStream_t stream) {
  if (count == 0) return;
  std::uint32_t const blocks = static_cast<std::uint32_t>((count + 255) / 256);
  hipLaunchKernelGGL(records_to_tracks_kernel, dim3(blocks), dim3(256), 0, stream, records,
                     count, tracks, stride);
  C8MI_HIP(hipGetLastError());
}

}  // namespace c8mi::radio
