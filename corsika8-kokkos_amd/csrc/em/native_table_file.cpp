// Format-v7 table sets on disk and their content hash. The payload of a table
// file is the canonical encoding the reference hashes
// (calculate_proposal_native_hash, src/kokkos/em/tables/ProposalNativeTable.cpp:
// 570-666: every integer little endian, every double by its bit pattern,
// strings and arrays with a u64 length, fields in the reference's order), so
// the SHA-256 stored after it is the table's content hash on both sides.
#include "em/native_table_file.hpp"

#include <fcntl.h>
#include <unistd.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>

#include "em/device_tables.hpp"

namespace c8mi::em {

namespace {

constexpr char kMagic[8] = {'C', '8', 'M', 'I', 'N', 'T', '0', '7'};

// ---- SHA-256 (FIPS 180-4) ----

constexpr std::uint32_t kRound[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

inline std::uint32_t rotr(std::uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

void compress(std::uint32_t h[8], std::uint8_t const block[64]) {
  std::uint32_t w[64];
  for (int t = 0; t < 16; ++t)
    w[t] = (std::uint32_t{block[4 * t]} << 24) | (std::uint32_t{block[4 * t + 1]} << 16) |
           (std::uint32_t{block[4 * t + 2]} << 8) | std::uint32_t{block[4 * t + 3]};
  for (int t = 16; t < 64; ++t) {
    std::uint32_t const s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
    std::uint32_t const s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
    w[t] = w[t - 16] + s0 + w[t - 7] + s1;
  }
  std::uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], k = h[7];
  for (int t = 0; t < 64; ++t) {
    std::uint32_t const t1 = k + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + kRound[t] + w[t];
    std::uint32_t const t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
    k = g;
    g = f;
    f = e;
    e = d + t1;
    d = c;
    c = b;
    b = a;
    a = t1 + t2;
  }
  h[0] += a;
  h[1] += b;
  h[2] += c;
  h[3] += d;
  h[4] += e;
  h[5] += f;
  h[6] += g;
  h[7] += k;
}

// ---- canonical encoding ----

class Encoder {
 public:
  void u32(std::uint32_t v) {
    for (int b = 0; b < 4; ++b) out_.push_back(static_cast<std::uint8_t>(v >> (8 * b)));
  }
  void i32(std::int32_t v) { u32(static_cast<std::uint32_t>(v)); }
  void u64(std::uint64_t v) {
    for (int b = 0; b < 8; ++b) out_.push_back(static_cast<std::uint8_t>(v >> (8 * b)));
  }
  void f64(double v) {
    std::uint64_t bits;
    std::memcpy(&bits, &v, sizeof bits);
    u64(bits);
  }
  void text(char const* s, std::size_t capacity) {
    std::size_t const n = strnlen(s, capacity);
    if (n == capacity) throw std::invalid_argument("native table: unterminated version string");
    u64(n);
    out_.insert(out_.end(), s, s + n);
  }
  void axis(c8mi_native_axis const& a) {
    u32(a.type);
    u32(a.nodes);
    f64(a.low);
    f64(a.high);
    f64(a.step);
  }
  void cubic(c8mi_native_cubic const& c) {
    axis(c.axis);
    axis(c.value_transform);
    u64(c.coefficient_offset);
    u32(c.coefficient_count);
    u32(c.has_value_transform);
  }
  void doubles(double const* v, std::uint64_t n) {
    u64(n);
    for (std::uint64_t i = 0; i < n; ++i) f64(v[i]);
  }
  std::vector<std::uint8_t>& bytes() { return out_; }

 private:
  std::vector<std::uint8_t> out_;
};

class Decoder {
 public:
  Decoder(std::uint8_t const* data, std::size_t size) : p_(data), end_(data + size) {}
  std::uint32_t u32() {
    need(4);
    std::uint32_t v = 0;
    for (int b = 0; b < 4; ++b) v |= std::uint32_t{p_[b]} << (8 * b);
    p_ += 4;
    return v;
  }
  std::int32_t i32() { return static_cast<std::int32_t>(u32()); }
  std::uint64_t u64() {
    need(8);
    std::uint64_t v = 0;
    for (int b = 0; b < 8; ++b) v |= std::uint64_t{p_[b]} << (8 * b);
    p_ += 8;
    return v;
  }
  double f64() {
    std::uint64_t const bits = u64();
    double v;
    std::memcpy(&v, &bits, sizeof v);
    return v;
  }
  void text(char* out, std::size_t capacity) {
    std::uint64_t const n = u64();
    if (n >= capacity) throw std::invalid_argument("native table file: version string too long");
    need(n);
    std::memcpy(out, p_, n);
    out[n] = '\0';
    p_ += n;
  }
  c8mi_native_axis axis() {
    c8mi_native_axis a{};
    a.type = u32();
    a.nodes = u32();
    a.low = f64();
    a.high = f64();
    a.step = f64();
    return a;
  }
  c8mi_native_cubic cubic() {
    c8mi_native_cubic c{};
    c.axis = axis();
    c.value_transform = axis();
    c.coefficient_offset = u64();
    c.coefficient_count = u32();
    c.has_value_transform = u32();
    return c;
  }
  // A sequence length, bounded by what the remaining bytes can hold.
  std::uint64_t count(std::size_t bytes_each) {
    std::uint64_t const n = u64();
    if (n > static_cast<std::uint64_t>(end_ - p_) / bytes_each)
      throw std::invalid_argument("native table file: truncated sequence");
    return n;
  }
  std::vector<double> doubles() {
    std::uint64_t const n = count(8);
    std::vector<double> v(n);
    for (auto& x : v) x = f64();
    return v;
  }
  bool done() const { return p_ == end_; }

 private:
  void need(std::size_t n) {
    if (static_cast<std::size_t>(end_ - p_) < n) throw std::invalid_argument("native table file: truncated");
  }
  std::uint8_t const* p_;
  std::uint8_t const* end_;
};

std::vector<std::uint8_t> encode(c8mi_native_table_set const& t, c8mi_native_table_meta const& m) {
  if (m.interaction_identity_count && !m.interaction_identities)
    throw std::invalid_argument("native table: null interaction identities");
  Encoder e;
  e.u32(m.format_version);
  e.text(m.proposal_version, sizeof m.proposal_version);
  e.text(m.cubic_interpolation_version, sizeof m.cubic_interpolation_version);
  c8mi_native_constants const& c = t.constants;
  for (double v : {c.electron_mass_mev, c.charged_pion_mass_mev, c.fine_structure_constant,
                   c.classical_electron_radius_cm, c.pi, c.sqrt_e, c.muon_mass_mev})
    e.f64(v);
  e.f64(m.stochastic_cut_mev);
  e.u64(m.interaction_identity_count);
  for (std::uint64_t i = 0; i < m.interaction_identity_count; ++i) {
    auto const& id = m.interaction_identities[i];
    e.i32(id.pdg_id);
    e.u32(id.reserved);
    e.u64(id.medium_hash);
    e.u64(id.interaction_hash);
  }
  e.u64(t.dndx_column_count);
  for (std::uint32_t i = 0; i < t.dndx_column_count; ++i) {
    auto const& d = t.dndx_columns[i];
    e.i32(d.pdg_id);
    e.i32(d.process_id);
    e.u64(d.medium_hash);
    e.u64(d.component_hash);
    e.u64(d.cross_section_hash);
    e.u64(d.table_hash);
    e.u32(d.loss_transform);
    e.u32(d.kinematic_model);
    e.u32(d.rate_model);
    e.axis(d.spline.energy_axis);
    e.axis(d.spline.loss_axis);
    e.u32(d.spline.rows);
    e.u32(d.spline.columns);
    e.u64(d.spline.coefficient_offset);
    e.u64(d.spline.polynomial_coefficient_offset);
    for (double v : {d.component_weight, d.particle_mass_mev, d.component_nuclear_charge, d.component_atomic_mass,
                     d.component_average_nucleon_mass_mev, d.direct_rate_factor, d.lower_energy_limit_mev,
                     d.absolute_cut_mev, d.relative_cut, d.medium_mean_excitation_energy_mev})
      e.f64(v);
    e.u32(d.selection_ordinal);
    e.u32(d.reserved);
  }
  e.u64(t.total_rate_column_count);
  for (std::uint32_t i = 0; i < t.total_rate_column_count; ++i) {
    auto const& r = t.total_rate_columns[i];
    e.i32(r.pdg_id);
    e.u32(r.reserved);
    e.u64(r.medium_hash);
    e.u64(r.table_hash);
    e.cubic(r.spline);
    e.f64(r.lower_energy_limit_mev);
  }
  e.u64(t.dndx_column_count);
  for (std::uint32_t i = 0; i < t.dndx_column_count; ++i) e.u32(t.selection_column_indices[i]);
  e.u64(t.dedx_column_count);
  for (std::uint32_t i = 0; i < t.dedx_column_count; ++i) {
    auto const& d = t.dedx_columns[i];
    e.i32(d.pdg_id);
    e.i32(d.process_id);
    e.u64(d.medium_hash);
    e.u64(d.cross_section_hash);
    e.u64(d.component_index);
    e.u64(d.table_hash);
    e.cubic(d.spline);
    e.f64(d.component_weight);
    e.f64(d.lower_energy_limit_mev);
    e.u32(d.accumulation_ordinal);
    e.u32(d.reserved);
  }
  e.u64(t.dedx_column_count);
  for (std::uint32_t i = 0; i < t.dedx_column_count; ++i) e.u32(t.dedx_accumulation_indices[i]);
  e.u64(t.utility_column_count);
  for (std::uint32_t i = 0; i < t.utility_column_count; ++i) {
    auto const& u = t.utility_columns[i];
    e.i32(u.pdg_id);
    e.u64(u.medium_hash);
    e.u64(u.table_hash);
    e.cubic(u.spline);
    e.f64(u.particle_mass_mev);
    e.f64(u.lower_energy_limit_mev);
    e.u32(u.reverse);
  }
  e.doubles(t.bicubic_values, t.bicubic_coefficient_count);
  e.doubles(t.bicubic_derivative_energy, t.bicubic_coefficient_count);
  e.doubles(t.bicubic_derivative_loss, t.bicubic_coefficient_count);
  e.doubles(t.bicubic_mixed_derivative, t.bicubic_coefficient_count);
  e.doubles(t.bicubic_polynomial_coefficients, t.bicubic_polynomial_coefficient_count);
  e.doubles(t.cubic_values, t.cubic_coefficient_count);
  e.doubles(t.cubic_node_derivatives, t.cubic_coefficient_count);
  return std::move(e.bytes());
}

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

std::vector<std::uint8_t> read_file(char const* path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error(std::string("cannot open the native table file ") + path);
  std::vector<std::uint8_t> bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  if (in.bad()) throw std::runtime_error(std::string("cannot read the native table file ") + path);
  return bytes;
}

}  // namespace

Digest sha256(std::uint8_t const* data, std::size_t size) {
  std::uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                        0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
  std::size_t full = size / 64;
  for (std::size_t i = 0; i < full; ++i) compress(h, data + 64 * i);
  std::uint8_t tail[128] = {};
  std::size_t const rest = size - 64 * full;
  if (rest) std::memcpy(tail, data + 64 * full, rest);
  tail[rest] = 0x80;
  std::size_t const blocks = rest + 9 > 64 ? 2 : 1;
  std::uint64_t const bits = static_cast<std::uint64_t>(size) * 8;
  for (int b = 0; b < 8; ++b) tail[64 * blocks - 1 - b] = static_cast<std::uint8_t>(bits >> (8 * b));
  for (std::size_t i = 0; i < blocks; ++i) compress(h, tail + 64 * i);
  Digest out;
  for (int i = 0; i < 8; ++i)
    for (int b = 0; b < 4; ++b) out[4 * i + b] = static_cast<std::uint8_t>(h[i] >> (24 - 8 * b));
  return out;
}

Digest native_table_hash(c8mi_native_table_set const& set, c8mi_native_table_meta const& meta) {
  std::vector<std::uint8_t> const payload = encode(set, meta);
  return sha256(payload.data(), payload.size());
}

TableFile::TableFile(std::vector<std::uint8_t> const& file) {
  if (file.size() < sizeof kMagic + 8 + 32 || std::memcmp(file.data(), kMagic, sizeof kMagic) != 0)
    throw std::invalid_argument("native table file: not a C8MINT07 file");
  std::uint64_t length = 0;
  for (int b = 0; b < 8; ++b) length |= std::uint64_t{file[sizeof kMagic + b]} << (8 * b);
  std::size_t const header = sizeof kMagic + 8;
  if (length != file.size() - header - 32) throw std::invalid_argument("native table file: length mismatch");
  std::uint8_t const* payload = file.data() + header;
  digest_ = sha256(payload, length);
  if (std::memcmp(digest_.data(), payload + length, 32) != 0)
    throw std::invalid_argument("native table file: content hash mismatch");
  Decoder d(payload, length);
  meta_ = c8mi_native_table_meta{};
  meta_.format_version = d.u32();
  if (meta_.format_version != C8MI_NATIVE_TABLE_FORMAT_VERSION)
    throw std::invalid_argument("native table file: unsupported format version");
  d.text(meta_.proposal_version, sizeof meta_.proposal_version);
  d.text(meta_.cubic_interpolation_version, sizeof meta_.cubic_interpolation_version);
  c8mi_native_constants& c = set_.constants;
  for (double* v : {&c.electron_mass_mev, &c.charged_pion_mass_mev, &c.fine_structure_constant,
                    &c.classical_electron_radius_cm, &c.pi, &c.sqrt_e, &c.muon_mass_mev})
    *v = d.f64();
  meta_.stochastic_cut_mev = d.f64();
  identities_.resize(d.count(24));
  for (auto& id : identities_) {
    id.pdg_id = d.i32();
    id.reserved = d.u32();
    id.medium_hash = d.u64();
    id.interaction_hash = d.u64();
  }
  dndx_.resize(d.count(4));
  for (auto& x : dndx_) {
    x = c8mi_native_dndx_column{};
    x.pdg_id = d.i32();
    x.process_id = d.i32();
    x.medium_hash = d.u64();
    x.component_hash = d.u64();
    x.cross_section_hash = d.u64();
    x.table_hash = d.u64();
    x.loss_transform = d.u32();
    x.kinematic_model = d.u32();
    x.rate_model = d.u32();
    x.spline.energy_axis = d.axis();
    x.spline.loss_axis = d.axis();
    x.spline.rows = d.u32();
    x.spline.columns = d.u32();
    x.spline.coefficient_offset = d.u64();
    x.spline.polynomial_coefficient_offset = d.u64();
    for (double* v : {&x.component_weight, &x.particle_mass_mev, &x.component_nuclear_charge,
                      &x.component_atomic_mass, &x.component_average_nucleon_mass_mev, &x.direct_rate_factor,
                      &x.lower_energy_limit_mev, &x.absolute_cut_mev, &x.relative_cut,
                      &x.medium_mean_excitation_energy_mev})
      *v = d.f64();
    x.selection_ordinal = d.u32();
    x.reserved = d.u32();
  }
  totals_.resize(d.count(4));
  for (auto& x : totals_) {
    x = c8mi_native_total_rate_column{};
    x.pdg_id = d.i32();
    x.reserved = d.u32();
    x.medium_hash = d.u64();
    x.table_hash = d.u64();
    x.spline = d.cubic();
    x.lower_energy_limit_mev = d.f64();
  }
  selection_.resize(d.count(4));
  for (auto& v : selection_) v = d.u32();
  dedx_.resize(d.count(4));
  for (auto& x : dedx_) {
    x = c8mi_native_dedx_column{};
    x.pdg_id = d.i32();
    x.process_id = d.i32();
    x.medium_hash = d.u64();
    x.cross_section_hash = d.u64();
    x.component_index = d.u64();
    x.table_hash = d.u64();
    x.spline = d.cubic();
    x.component_weight = d.f64();
    x.lower_energy_limit_mev = d.f64();
    x.accumulation_ordinal = d.u32();
    x.reserved = d.u32();
  }
  accumulation_.resize(d.count(4));
  for (auto& v : accumulation_) v = d.u32();
  utility_.resize(d.count(4));
  for (auto& x : utility_) {
    x = c8mi_native_utility_column{};
    x.pdg_id = d.i32();
    x.medium_hash = d.u64();
    x.table_hash = d.u64();
    x.spline = d.cubic();
    x.particle_mass_mev = d.f64();
    x.lower_energy_limit_mev = d.f64();
    x.reverse = d.u32();
  }
  for (auto* v : {&bicubic_[0], &bicubic_[1], &bicubic_[2], &bicubic_[3], &polynomial_, &cubic_[0], &cubic_[1]})
    *v = d.doubles();
  if (!d.done()) throw std::invalid_argument("native table file: trailing bytes in the payload");
  // the set keeps one count per family of arrays
  if (selection_.size() != dndx_.size() || accumulation_.size() != dedx_.size())
    throw std::invalid_argument("native table file: index arrays do not match their columns");
  for (int k = 1; k < 4; ++k)
    if (bicubic_[k].size() != bicubic_[0].size())
      throw std::invalid_argument("native table file: bicubic arrays of different lengths");
  if (cubic_[1].size() != cubic_[0].size())
    throw std::invalid_argument("native table file: cubic arrays of different lengths");
  if (dndx_.size() > 0xffffffffu || totals_.size() > 0xffffffffu || dedx_.size() > 0xffffffffu ||
      utility_.size() > 0xffffffffu)
    throw std::invalid_argument("native table file: too many columns");
  auto ptr = [](auto const& v) { return v.empty() ? nullptr : v.data(); };
  set_.dndx_columns = ptr(dndx_);
  set_.dndx_column_count = static_cast<std::uint32_t>(dndx_.size());
  set_.total_rate_columns = ptr(totals_);
  set_.total_rate_column_count = static_cast<std::uint32_t>(totals_.size());
  set_.selection_column_indices = ptr(selection_);
  set_.dedx_columns = ptr(dedx_);
  set_.dedx_accumulation_indices = ptr(accumulation_);
  set_.dedx_column_count = static_cast<std::uint32_t>(dedx_.size());
  set_.utility_columns = ptr(utility_);
  set_.utility_column_count = static_cast<std::uint32_t>(utility_.size());
  set_.bicubic_values = ptr(bicubic_[0]);
  set_.bicubic_derivative_energy = ptr(bicubic_[1]);
  set_.bicubic_derivative_loss = ptr(bicubic_[2]);
  set_.bicubic_mixed_derivative = ptr(bicubic_[3]);
  set_.bicubic_coefficient_count = bicubic_[0].size();
  set_.bicubic_polynomial_coefficients = ptr(polynomial_);
  set_.bicubic_polynomial_coefficient_count = polynomial_.size();
  set_.cubic_values = ptr(cubic_[0]);
  set_.cubic_node_derivatives = ptr(cubic_[1]);
  set_.cubic_coefficient_count = cubic_[0].size();
  meta_.interaction_identities = ptr(identities_);
  meta_.interaction_identity_count = identities_.size();
  validate_table_bounds(set_);  // every descriptor within its arrays
}

std::unique_ptr<TableFile> read_native_table(char const* path) {
  return std::make_unique<TableFile>(read_file(path));
}

void write_native_table(char const* path, c8mi_native_table_set const& set, c8mi_native_table_meta const& meta) {
  validate_table_bounds(set);
  std::vector<std::uint8_t> const payload = encode(set, meta);
  Digest const digest = sha256(payload.data(), payload.size());
  std::uint8_t length[8];
  for (int b = 0; b < 8; ++b) length[b] = static_cast<std::uint8_t>(static_cast<std::uint64_t>(payload.size()) >> (8 * b));
  std::string const target(path);
  std::vector<char> name(target.begin(), target.end());
  char const suffix[] = ".tmp.XXXXXX";
  name.insert(name.end(), suffix, suffix + sizeof suffix);
  int const fd = ::mkstemp(name.data());
  if (fd < 0) throw std::runtime_error("cannot create a temporary file for the native table " + target);
  std::string const temporary(name.data());
  bool ok = write_all(fd, kMagic, sizeof kMagic) && write_all(fd, length, sizeof length) &&
            write_all(fd, payload.data(), payload.size()) && write_all(fd, digest.data(), digest.size()) &&
            ::fsync(fd) == 0;
  ok = (::close(fd) == 0) && ok;
  if (ok) {
    try {
      ok = TableFile(read_file(temporary.c_str())).digest() == digest;
    } catch (std::exception const&) {
      ok = false;
    }
  }
  if (ok) ok = ::rename(temporary.c_str(), target.c_str()) == 0;
  if (!ok) {
    ::unlink(temporary.c_str());
    throw std::runtime_error("cannot write the native table file " + target);
  }
  std::string const parent = target.find('/') == std::string::npos ? "." : target.substr(0, target.rfind('/') + 1);
  int const dir = ::open(parent.c_str(), O_RDONLY | O_DIRECTORY);
  if (dir >= 0) {
    (void)::fsync(dir);
    ::close(dir);
  }
}

}  // namespace c8mi::em
