// Host construction and cache I/O of the Molière interpolation table
// (c8mi_moliere_interpolation).
//
// Reference: src/kokkos/em/tables/MoliereInterpolation.cpp
//   build_function_polynomials :272-310  (Hermite form of a cubic spline per interval)
//   build_dimensionless_component_b :312-339, mixture_cumulative :341-360,
//   build_initial_guess_table :362-436, make_moliere_interpolation_table :440-460,
//   read/write_interpolation_cache :204-270 (FNV-1a checksum :186-196),
// and the analytical functions f1_m / f2_m / their integrals and
// evaluate_interpolation of corsika/detail/kokkos/em/models/MoliereScattering.inl.
//
// The reference's splines come from the CubicInterpolation library (cardinal
// cubic splines on a linear axis), which is not available here. This file
// interpolates with natural end conditions (zero second derivative at x = 0 and
// x = 20): the node values are the functions' own, the node derivatives those
// of the natural spline.
#include "em/moliere_table.hpp"

#include <array>
#include <cerrno>
#include <memory>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <fcntl.h>
#include <unistd.h>

namespace c8mi::em {
namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr int kIntervals = C8MI_MOLIERE_INTERVALS;
constexpr int kNodes = kIntervals + 1;
constexpr double kLower = 0.;
constexpr double kUpper = 20.;
constexpr int kBNodes = C8MI_MOLIERE_GUESS_B_NODES;
constexpr int kBetaNodes = C8MI_MOLIERE_GUESS_BETA2_NODES;
constexpr int kCoordinateNodes = C8MI_MOLIERE_GUESS_COORDINATE_NODES;
constexpr double kBLower = 4.5, kBUpper = 64.;
constexpr double kBetaLower = 0.25, kBetaUpper = 1.;
constexpr double kCoordinateLower = 0., kCoordinateUpper = 5.;
constexpr double kMinimumB = 4.5;
constexpr double kMinimumBRhs = 2.9959226032237258;

// ---- analytical functions (MoliereScattering.inl f1_m, f2_m and integrals) ----

double f1(c8mi_moliere_snapshot const& m, double x) {
  if (x > 12.) return 0.5 * std::sqrt(kPi) / (std::pow(x, 1.5) * std::pow(1. - 4.5 / x, 2. / 3.));
  double sum = m.c1[69];
  for (int i = 68; i >= 0; --i) sum = sum * x + m.c1[i];
  return sum;
}

double f2(c8mi_moliere_snapshot const& m, double x) {
  if (x > 4.25 * 4.25) {
    if (x <= 6.5 * 6.5)
      return 0.00013567765224589459194769192063035 * x - 0.0022635502525409950842771866774683 * std::sqrt(x) +
             0.0098037758070269476889935233998585;
    double sum = 0.;
    for (int i = 2; i < 13; ++i) sum += m.c2_large[i] * (0.5 * std::log(x) + m.s2_large[i]) * std::pow(x, -(i + 0.5));
    return sum;
  }
  double sum = m.c2[69];
  for (int i = 68; i >= 0; --i) sum = sum * x + m.c2[i];
  return sum;
}

double integral_f1(c8mi_moliere_snapshot const& m, double x) {
  if (x > 12.) {
    double sum = m.c1_integral_large[14];
    for (int i = 13; i >= 0; --i) sum = m.c1_integral_large[i] + sum / x;
    return sum;
  }
  double sum = m.c1[69] / (2. * 69. + 1.);
  for (int i = 68; i >= 0; --i) sum = sum * x + m.c1[i] / (2. * i + 1.);
  return sum * std::sqrt(x > 0. ? x : 0.);
}

double integral_f2(c8mi_moliere_snapshot const& m, double x) {
  if (x > 4.25 * 4.25) {
    if (x <= 6.5 * 6.5)
      return -0.00026360133958801203364619158975302 * x + 0.0039965027465457608410459577896745 * std::sqrt(x) -
             0.016305842044996649714549974419242;
    double sum = 0.;
    for (int i = 2; i < 13; ++i)
      sum += -0.5 * m.c2_large[i] / i * (0.5 / i + 0.5 * std::log(x) + m.s2_large[i]) * std::pow(x, -i);
    return sum;
  }
  double sum = m.c2[69] / (2. * 69. + 1.);
  for (int i = 68; i >= 0; --i) sum = sum * x + m.c2[i] / (2. * i + 1.);
  return sum * std::sqrt(x > 0. ? x : 0.);
}

double solve_b(double rhs) {
  if (!std::isfinite(rhs) || rhs < kMinimumBRhs) return 0.;
  double const l = std::log(rhs), ir = 1. / rhs, ir2 = ir * ir, l2 = l * l, l3 = l2 * l;
  double b = rhs + l + l * ir + l * (2. - l) * 0.5 * ir2 + l * (6. - 9. * l + 2. * l2) / 6. * ir2 * ir +
             l * (12. - 36. * l + 22. * l2 - 3. * l3) / 12. * ir2 * ir2;
  if (!std::isfinite(b) || !(b > 1.)) return 0.;
  double const f = b - std::log(b) - rhs, ib = 1. / b, df = 1. - ib;
  double const den = 2. * df * df - f * ib * ib;
  if (!std::isfinite(den) || den == 0.) return 0.;
  b -= 2. * f * df / den;
  return std::isfinite(b) && b >= kMinimumB ? b : 0.;
}

// ---- splines ----

// Hermite coefficients of every interval of the natural cubic spline through
// (x_k, f(x_k)), x_k = k h on [0, 20].
template <class F>
void spline_polynomials(F const& function, c8mi_moliere_cubic* out) {
  double const h = (kUpper - kLower) / static_cast<double>(kIntervals);
  std::array<double, kNodes> y{}, second{}, slope{};
  for (int k = 0; k < kNodes; ++k) y[k] = function(kLower + h * static_cast<double>(k));
  // tridiagonal system for the interior second derivatives (natural ends: 0)
  std::array<double, kNodes> diag{}, rhs{};
  for (int k = 1; k < kNodes - 1; ++k) {
    diag[k] = 4.;
    rhs[k] = 6. * (y[k + 1] - 2. * y[k] + y[k - 1]) / (h * h);
  }
  for (int k = 2; k < kNodes - 1; ++k) {  // forward elimination, unit off-diagonals
    double const w = 1. / diag[k - 1];
    diag[k] -= w;
    rhs[k] -= w * rhs[k - 1];
  }
  second[kNodes - 1] = 0.;
  for (int k = kNodes - 2; k >= 1; --k) second[k] = (rhs[k] - second[k + 1]) / diag[k];
  second[0] = 0.;
  for (int k = 0; k < kNodes - 1; ++k)
    slope[k] = (y[k + 1] - y[k]) / h - h * (2. * second[k] + second[k + 1]) / 6.;
  slope[kNodes - 1] = (y[kNodes - 1] - y[kNodes - 2]) / h + h * (second[kNodes - 2] + 2. * second[kNodes - 1]) / 6.;
  for (int k = 0; k < kIntervals; ++k) {
    double const y0 = y[k], y1 = y[k + 1];
    double const d0 = h * slope[k], d1 = h * slope[k + 1];
    c8mi_moliere_cubic& p = out[k];
    p.c[0] = y0;
    p.c[1] = d0;
    p.c[2] = 3. * (y1 - y0) - 2. * d0 - d1;
    p.c[3] = 2. * (y0 - y1) + d0 + d1;
    for (double v : p.c)
      if (!std::isfinite(v)) throw std::runtime_error("Moliere interpolation produced a non-finite coefficient");
  }
}

// evaluate_interpolation (MoliereScattering.inl:118-134).
double interpolate(c8mi_moliere_cubic const* p, double x) {
  double const inverse_step = static_cast<double>(kIntervals) / (kUpper - kLower);
  double coordinate = (x - kLower) * inverse_step;
  if (coordinate < 0.) coordinate = 0.;
  if (coordinate > static_cast<double>(kIntervals)) coordinate = static_cast<double>(kIntervals);
  auto interval = static_cast<std::uint32_t>(coordinate);
  if (interval >= static_cast<std::uint32_t>(kIntervals)) interval = kIntervals - 1;
  double const q = coordinate - static_cast<double>(interval);
  c8mi_moliere_cubic const& c = p[interval];
  return c.c[0] + q * (c.c[1] + q * (c.c[2] + q * c.c[3]));
}

// ---- initial-guess corrections ----

bool component_b(c8mi_moliere_snapshot const& m, double reference_b, double beta2, double* b) {
  if (!(reference_b >= kMinimumB) || !(beta2 > 0.) || !std::isfinite(reference_b) || !std::isfinite(beta2))
    return false;
  auto const& ref = m.components[m.maximum_weight_index];
  double const reference_screening = ref.chi0_squared_mev2 * (1.13 + ref.coulomb_correction / beta2);
  if (!(reference_screening > 0.) || !std::isfinite(reference_screening)) return false;
  double const reference_rhs = reference_b - std::log(reference_b);
  for (std::uint32_t i = 0; i < m.component_count; ++i) {
    double const screening =
        m.components[i].chi0_squared_mev2 * (1.13 + m.components[i].coulomb_correction / beta2);
    if (!(screening > 0.) || !std::isfinite(screening)) return false;
    b[i] = solve_b(reference_rhs + std::log(reference_screening / screening));
    if (!(b[i] > 0.) || !std::isfinite(b[i])) return false;
  }
  return true;
}

double mixture_cumulative(c8mi_moliere_snapshot const& m, c8mi_moliere_cubic const* polynomials,
                          double const* b, double reference_b, double coordinate) {
  double result = 0.;
  double const c2 = coordinate * coordinate;
  double const inverse_sqrt_pi = std::sqrt(1. / kPi);
  c8mi_moliere_cubic const* i1 = polynomials + 2 * kIntervals;
  c8mi_moliere_cubic const* i2 = polynomials + 3 * kIntervals;
  for (std::uint32_t i = 0; i < m.component_count; ++i) {
    double const inverse_b = 1. / b[i];
    double const x = c2 * reference_b * inverse_b;
    double const inverse_b2 = inverse_b * inverse_b;
    double const v1 = x <= kUpper ? interpolate(i1, x) : integral_f1(m, x);
    double const v2 = x <= kUpper ? interpolate(i2, x) : integral_f2(m, x);
    result += m.components[i].weight_zz *
              (0.5 * std::erf(std::sqrt(x)) + inverse_sqrt_pi * (v1 * inverse_b + v2 * inverse_b2));
  }
  return result * m.inverse_weight_zz_sum;
}

void initial_guess(c8mi_moliere_snapshot const& m, c8mi_moliere_cubic const* polynomials, double* out) {
  constexpr double b_step = (kBUpper - kBLower) / static_cast<double>(kBNodes - 1);
  constexpr double coordinate_step = (kCoordinateUpper - kCoordinateLower) / static_cast<double>(kCoordinateNodes - 1);
  constexpr double beta_step = (kBetaUpper - kBetaLower) / static_cast<double>(kBetaNodes - 1);
  double b[C8MI_MOLIERE_MAX_COMPONENTS];
  for (int bi = 0; bi < kBNodes; ++bi) {
    double const reference_b = kBLower + static_cast<double>(bi) * b_step;
    for (int ti = 0; ti < kBetaNodes; ++ti) {
      double const beta2 = kBetaLower + static_cast<double>(ti) * beta_step;
      bool const valid = component_b(m, reference_b, beta2, b);
      for (int ci = 0; ci < kCoordinateNodes; ++ci) {
        double const gaussian = kCoordinateLower + static_cast<double>(ci) * coordinate_step;
        double& slot = out[(bi * kBetaNodes + ti) * kCoordinateNodes + ci];
        if (ci == 0 || !valid) {
          slot = 0.;
          continue;
        }
        double const target = 0.5 * std::erf(gaussian);
        double lower = 0., upper = gaussian > 1. ? gaussian : 1.;
        double upper_cdf = mixture_cumulative(m, polynomials, b, reference_b, upper);
        for (int e = 0; e < 32 && std::isfinite(upper_cdf) && upper_cdf < target; ++e) {
          upper *= 2.;
          upper_cdf = mixture_cumulative(m, polynomials, b, reference_b, upper);
        }
        if (!std::isfinite(upper_cdf) || upper_cdf < target) {  // the Gaussian guess stays
          slot = 0.;
          continue;
        }
        for (int it = 0; it < 24; ++it) {
          double const mid = 0.5 * (lower + upper);
          double const cdf = mixture_cumulative(m, polynomials, b, reference_b, mid);
          if (!std::isfinite(cdf)) {
            lower = upper = gaussian;
            break;
          }
          if (cdf < target) lower = mid;
          else upper = mid;
        }
        double const delta = 0.5 * (lower + upper) - gaussian;
        slot = std::isfinite(delta) ? delta : 0.;
      }
    }
  }
}

// ---- cache file ----

// The reference's cache magic (its own CubicInterpolation-built tables), and the
// magic this build writes: its natural-spline tables differ from the
// reference's in the last digits, so a file written here must not pass the
// reference's reader (which checks the magic) as one of its own.
constexpr char kMagic[8] = {'C', '8', 'M', 'O', 'L', 'I', '0', '1'};
constexpr char kBuilderMagic[8] = {'C', '8', 'M', 'I', 'M', 'O', '0', '1'};
constexpr std::uint32_t kFormatVersion = 1;

struct CacheHeader {
  char magic[8];
  std::uint32_t format_version;
  std::uint32_t snapshot_bytes;
  std::uint64_t table_bytes;
  std::uint64_t table_checksum;
  c8mi_moliere_snapshot snapshot;
};

std::uint64_t fnv1a(void const* data, std::size_t size) {
  auto const* bytes = static_cast<unsigned char const*>(data);
  std::uint64_t h = 14695981039346656037ULL;
  for (std::size_t i = 0; i < size; ++i) {
    h ^= bytes[i];
    h *= 1099511628211ULL;
  }
  return h;
}

}  // namespace

bool valid_moliere_snapshot(c8mi_moliere_snapshot const& m) {
  auto pos = [](double v) { return std::isfinite(v) && v > 0.; };
  if (m.component_count == 0 || m.component_count > C8MI_MOLIERE_MAX_COMPONENTS ||
      m.maximum_weight_index >= m.component_count || !pos(m.particle_mass_mev) || !pos(m.electron_mass_mev) ||
      !pos(m.fine_structure_constant) || !pos(m.avogadro_per_mol) || !pos(m.hbar_mev_s) ||
      !pos(m.speed_of_light_cm_per_s) || !pos(m.z_squared_over_a_average) || !pos(m.inverse_weight_zz_sum) ||
      !std::isfinite(m.euler_mascheroni))
    return false;
  for (std::uint32_t k = 0; k < m.component_count; ++k) {
    c8mi_moliere_component const& c = m.components[k];
    if (!pos(c.nuclear_charge) || !pos(c.weight_zz) || !pos(c.chi0_squared_mev2) || !pos(c.coulomb_correction))
      return false;
  }
  for (int k = 0; k < 70; ++k)
    if (!std::isfinite(m.c1[k]) || !std::isfinite(m.c2[k])) return false;
  for (int k = 0; k < 13; ++k)
    if (!std::isfinite(m.c2_large[k]) || !std::isfinite(m.s2_large[k])) return false;
  for (int k = 0; k < 15; ++k)
    if (!std::isfinite(m.c1_integral_large[k])) return false;
  return true;
}

bool valid_moliere_table(c8mi_moliere_interpolation const& table) {
  for (auto const& p : table.polynomials)
    for (double v : p.c)
      if (!std::isfinite(v)) return false;
  for (double v : table.initial_guess_delta)
    if (!std::isfinite(v)) return false;
  return true;
}

void build_moliere_table(c8mi_moliere_snapshot const& m, c8mi_moliere_interpolation& out) {
  if (!valid_moliere_snapshot(m)) throw std::invalid_argument("cannot interpolate an invalid Moliere snapshot");
  std::memset(&out, 0, sizeof out);
  spline_polynomials([&](double x) { return f1(m, x); }, out.polynomials);
  spline_polynomials([&](double x) { return f2(m, x); }, out.polynomials + kIntervals);
  spline_polynomials([&](double x) { return integral_f1(m, x); }, out.polynomials + 2 * kIntervals);
  spline_polynomials([&](double x) { return integral_f2(m, x); }, out.polynomials + 3 * kIntervals);
  initial_guess(m, out.polynomials, out.initial_guess_delta);
}

bool read_moliere_table(char const* path, c8mi_moliere_snapshot const& snapshot, c8mi_moliere_interpolation& out) {
  std::ifstream in(path, std::ios::binary | std::ios::ate);
  if (!in) return false;
  auto const size = in.tellg();
  if (size < 0 || static_cast<std::uint64_t>(size) != sizeof(CacheHeader) + sizeof(c8mi_moliere_interpolation))
    return false;
  in.seekg(0);
  CacheHeader header{};
  in.read(reinterpret_cast<char*>(&header), sizeof header);
  in.read(reinterpret_cast<char*>(&out), sizeof out);
  bool const magic = std::memcmp(header.magic, kMagic, sizeof kMagic) == 0 ||
                     std::memcmp(header.magic, kBuilderMagic, sizeof kBuilderMagic) == 0;
  return static_cast<bool>(in) && magic && header.format_version == kFormatVersion &&
         header.snapshot_bytes == sizeof(c8mi_moliere_snapshot) &&
         header.table_bytes == sizeof(c8mi_moliere_interpolation) &&
         std::memcmp(&header.snapshot, &snapshot, sizeof snapshot) == 0 &&
         header.table_checksum == fnv1a(&out, sizeof out) && valid_moliere_table(out);
}

namespace {

bool write_all(int fd, void const* data, std::size_t size) {
  auto const* p = static_cast<char const*>(data);
  while (size) {
    ssize_t const n = ::write(fd, p, size);
    if (n < 0) {
      if (errno == EINTR) continue;
      return false;
    }
    p += n;
    size -= static_cast<std::size_t>(n);
  }
  return true;
}

}  // namespace

// write_interpolation_cache (MoliereInterpolation.cpp:233-270): the table goes
// to a private temporary file next to the target, is synced, read back and
// compared, and only then renamed over the target, so a reader never sees a
// partial file. Written with this build's magic (see kBuilderMagic).
void write_moliere_table(char const* path, c8mi_moliere_snapshot const& snapshot,
                         c8mi_moliere_interpolation const& table) {
  CacheHeader header{};
  std::memcpy(header.magic, kBuilderMagic, sizeof kBuilderMagic);
  header.format_version = kFormatVersion;
  header.snapshot_bytes = sizeof(c8mi_moliere_snapshot);
  header.table_bytes = sizeof(c8mi_moliere_interpolation);
  header.table_checksum = fnv1a(&table, sizeof table);
  header.snapshot = snapshot;
  std::string const target(path);
  std::vector<char> name(target.begin(), target.end());
  char const suffix[] = ".tmp.XXXXXX";
  name.insert(name.end(), suffix, suffix + sizeof suffix);  // with the terminator
  int const fd = ::mkstemp(name.data());
  if (fd < 0) throw std::runtime_error("cannot create a temporary file for the Moliere cache " + target);
  std::string const temporary(name.data());
  bool ok = write_all(fd, &header, sizeof header) && write_all(fd, &table, sizeof table) && ::fsync(fd) == 0;
  ok = (::close(fd) == 0) && ok;
  if (ok) {
    auto verified = std::make_unique<c8mi_moliere_interpolation>();
    ok = read_moliere_table(temporary.c_str(), snapshot, *verified) &&
         std::memcmp(verified.get(), &table, sizeof table) == 0;
  }
  if (ok) ok = ::rename(temporary.c_str(), target.c_str()) == 0;
  if (!ok) {
    ::unlink(temporary.c_str());
    throw std::runtime_error("cannot write the Moliere cache file " + target);
  }
  std::string const parent = target.find('/') == std::string::npos ? "." : target.substr(0, target.rfind('/') + 1);
  int const dir = ::open(parent.c_str(), O_RDONLY | O_DIRECTORY);
  if (dir >= 0) {
    (void)::fsync(dir);
    ::close(dir);
  }
}

}  // namespace c8mi::em
