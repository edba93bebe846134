#include "em/em_backend.hpp"

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "em/em_front.hpp"

namespace c8mi::em {

namespace {

void require(bool ok, char const* message) {
  if (!ok) throw std::invalid_argument(std::string("em: ") + message);
}

template <class T>
void upload(DeviceBuffer<T>& buffer, T const* host, std::size_t count) {
  buffer.allocate(count);
  if (count) C8MI_HIP(hipMemcpy(buffer.data(), host, count * sizeof(T), hipMemcpyHostToDevice));
}

void validate_environment(c8mi_environment const& e) {
  require(e.node_count > 0 && e.nodes != nullptr, "environment has no nodes");
  require(e.model_count > 0 && e.models != nullptr, "environment has no medium models");
  require(e.child_index_count == 0 || e.children != nullptr, "environment child list");
  require(std::isfinite(e.maximum_magnetic_deflection_rad) && e.maximum_magnetic_deflection_rad > 0. &&
              e.maximum_magnetic_deflection_rad < 0.5 * 3.141592653589793,
          "maximum magnetic deflection must lie in (0, pi/2)");
  for (std::uint32_t i = 0; i < e.node_count; ++i) {
    c8mi_env_node const& n = e.nodes[i];
    require(n.shape_kind == C8MI_SHAPE_UNIVERSE || n.shape_kind == C8MI_SHAPE_SPHERE,
            "unsupported environment shape (universe and spheres only)");
    require(n.shape_kind != C8MI_SHAPE_SPHERE || (std::isfinite(n.radius_m) && n.radius_m > 0.),
            "sphere radius");
    require(n.parent >= -1 && n.parent < static_cast<std::int32_t>(e.node_count), "node parent");
    require(n.model >= -1 && n.model < static_cast<std::int32_t>(e.model_count), "node model");
    require(n.child_begin <= e.child_index_count && n.child_count <= e.child_index_count - n.child_begin,
            "node child range");
    for (std::uint32_t k = 0; k < n.child_count; ++k) {
      std::int32_t const c = e.children[n.child_begin + k];
      require(c >= 0 && c < static_cast<std::int32_t>(e.node_count), "child index");
    }
  }
  for (std::uint32_t i = 0; i < e.model_count; ++i) {
    c8mi_env_model const& m = e.models[i];
    require(m.density_kind == C8MI_DENSITY_HOMOGENEOUS ||
                m.density_kind == C8MI_DENSITY_SLIDING_PLANAR_EXPONENTIAL ||
                m.density_kind == C8MI_DENSITY_FLAT_EXPONENTIAL,
            "unsupported density model");
    require(std::isfinite(m.reference_density_g_per_cm3) && m.reference_density_g_per_cm3 > 0.,
            "reference density");
    require(m.magnetic_field_t[0] == 0. && m.magnetic_field_t[1] == 0. && m.magnetic_field_t[2] == 0.,
            "magnetic fields are not transported by this build (the field must be zero)");
  }
  if (e.observation_enabled) {
    double const n2 = e.observation_normal[0] * e.observation_normal[0] +
                      e.observation_normal[1] * e.observation_normal[1] +
                      e.observation_normal[2] * e.observation_normal[2];
    require(std::fabs(n2 - 1.) <= 1.e-12, "observation plane normal must be a unit vector");
  }
}

void validate_config(c8mi_em_config const& c) {
  require(c.queue_capacity > 0, "queue capacity");
  require(std::isfinite(c.em_transport_cut_mev) && c.em_transport_cut_mev > 0., "EM transport cut");
  require(c.profile_bins <= kMaxProfileBins, "profile bins above the supported maximum (1024)");
  require(c.profile_bins == 0 || (std::isfinite(c.profile_bin_width_g_per_cm2) && c.profile_bin_width_g_per_cm2 > 0.),
          "profile bin width");
  require(std::isfinite(c.profile_weight_scale) && c.profile_weight_scale > 0., "profile weight scale");
  require(std::isfinite(c.profile_energy_scale) && c.profile_energy_scale > 0., "profile energy scale");
  require(c.axis_grammage_g_per_cm2 != nullptr && c.axis_support_count > 0, "profile projection table");
  require(std::isfinite(c.axis_step_m) && c.axis_step_m > 0., "projection step");
  require(c.profile_crossing_mode <= C8MI_PROFILE_CROSSING_FORWARD, "profile crossing mode");
}

}  // namespace

QueueStorage::QueueStorage(std::uint64_t capacity) : capacity_(capacity) {
  // 4-byte fields first, then 8-byte fields, each array 256-byte aligned
  std::size_t const n = capacity;
  auto round = [](std::size_t b) { return (b + 255u) & ~std::size_t{255u}; };
  std::size_t const b4 = round(4 * n), b8 = round(8 * n);
  storage_.allocate(5 * b4 + 12 * b8);
  unsigned char* p = storage_.data();
  auto take4 = [&](auto*& field) {
    field = reinterpret_cast<std::remove_reference_t<decltype(field)>>(p);
    p += b4;
  };
  auto take8 = [&](auto*& field) {
    field = reinterpret_cast<std::remove_reference_t<decltype(field)>>(p);
    p += b8;
  };
  take4(view_.pid);
  take4(view_.medium);
  take4(view_.generation);
  take4(view_.reserved);
  take4(view_.node);
  take8(view_.energy);
  take8(view_.x);
  take8(view_.y);
  take8(view_.z);
  take8(view_.dx);
  take8(view_.dy);
  take8(view_.dz);
  take8(view_.time);
  take8(view_.weight);
  take8(view_.history);
  take8(view_.parent);
  take8(view_.step);
}

EmBackend::EmBackend(c8mi_em_config const& config, c8mi_native_table_set const& tables,
                     c8mi_environment const& environment, hipStream_t stream)
    : config_(config), tables_(tables), host_totals_(1) {
  validate_config(config);
  validate_environment(environment);
  if (stream) {
    stream_ = stream;
  } else {
    C8MI_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    own_stream_ = true;
  }
  upload(nodes_, environment.nodes, environment.node_count);
  upload(children_, environment.children, environment.child_index_count);
  upload(models_, environment.models, environment.model_count);
  upload(axis_, config.axis_grammage_g_per_cm2, config.axis_support_count);
  counters_.allocate(kCounters);
  profile_.allocate(4u * (config.profile_bins ? config.profile_bins : 1u));
  observations_.allocate(config.observation_capacity ? config.observation_capacity : 1);
  fallbacks_.allocate(config.fallback_capacity ? config.fallback_capacity : 1);
  record_counts_.allocate(2);
  for (auto& species : queues_)
    for (auto& q : species) q = std::make_unique<QueueStorage>(config.queue_capacity);

  EmContext& h = host_context_;
  h.tables = tables_.view();
  h.nodes = nodes_.data();
  h.children = children_.data();
  h.models = models_.data();
  h.node_count = environment.node_count;
  h.model_count = environment.model_count;
  h.observation_enabled = environment.observation_enabled;
  for (int a = 0; a < 3; ++a) {
    h.observation_point[a] = environment.observation_point_m[a];
    h.observation_normal[a] = environment.observation_normal[a];
    h.axis_start[a] = config.axis_start_m[a];
    h.axis_direction[a] = config.axis_direction[a];
  }
  h.maximum_magnetic_deflection = environment.maximum_magnetic_deflection_rad;
  h.seed = config.random_seed;
  h.shower = 0;
  h.em_cut_mev = config.em_transport_cut_mev;
  h.muon_cut_mev = config.muon_transport_cut_mev;
  h.pair_lpm = config.photon_pair_lpm;
  h.brems_lpm = config.brems_lpm;
  h.thinning = config.thinning;
  h.profile_bins = config.profile_bins;
  h.crossing_mode = config.profile_crossing_mode;
  h.bin_width = config.profile_bin_width_g_per_cm2;
  h.energy_loss_threshold = config.profile_energy_loss_threshold_g_per_cm2;
  h.weight_scale = config.profile_weight_scale;
  h.energy_scale = config.profile_energy_scale;
  h.axis_step = config.axis_step_m;
  h.axis_grammage = axis_.data();
  h.axis_support_count = config.axis_support_count;
  h.observations = observations_.data();
  h.observation_capacity = config.observation_capacity;
  h.fallbacks = fallbacks_.data();
  h.fallback_capacity = config.fallback_capacity;
  h.record_counts = record_counts_.data();
  h.counters = counters_.data();
  h.profile = profile_.data();
  context_.allocate(1);
  C8MI_HIP(hipMemcpy(context_.data(), &h, sizeof(EmContext), hipMemcpyHostToDevice));
  for (auto& e : events_) C8MI_HIP(hipEventCreate(&e));
  begin_shower(0, 1);
}

EmBackend::~EmBackend() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& e : events_)
    if (e) (void)hipEventDestroy(e);
  if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

void EmBackend::begin_shower(std::uint64_t shower_id, std::uint64_t first_history_id) {
  host_context_.shower = shower_id;
  C8MI_HIP(hipMemcpyAsync(context_.data(), &host_context_, sizeof(EmContext), hipMemcpyHostToDevice, stream_));
  C8MI_HIP(hipMemsetAsync(counters_.data(), 0, counters_.bytes(), stream_));
  C8MI_HIP(hipMemsetAsync(profile_.data(), 0, profile_.bytes(), stream_));
  C8MI_HIP(hipMemsetAsync(record_counts_.data(), 0, record_counts_.bytes(), stream_));
  C8MI_HIP(hipStreamSynchronize(stream_));
  count_[0] = count_[1] = 0;
  current_[0] = current_[1] = 0;
  next_history_ = first_history_id;
  injected_ = 0;
}

bool EmBackend::can_transport(c8mi_particle_state const& p) const {
  if (!(p.pid == 22 || p.pid == 11 || p.pid == -11)) return false;
  if (!std::isfinite(p.energy_gev) || !(p.energy_gev > 0.) || !std::isfinite(p.weight) || p.weight < 0.)
    return false;
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(p.position_m[a]) || !std::isfinite(p.direction[a])) return false;
  return true;
}

void EmBackend::enqueue(c8mi_particle_state const* particles, std::uint64_t count) {
  if (count == 0) return;
  if (!particles) throw std::invalid_argument("em: null particle array");
  std::vector<c8mi_particle_state> split[2];
  for (std::uint64_t i = 0; i < count; ++i) {
    c8mi_particle_state const& p = particles[i];
    if (!can_transport(p))
      throw std::invalid_argument("em: particle " + std::to_string(i) + " (pid " + std::to_string(p.pid) +
                                  ") cannot be transported by the device path");
    split[p.pid == 22 ? 0 : 1].push_back(p);
  }
  for (int s = 0; s < 2; ++s) {
    auto const& v = split[s];
    if (v.empty()) continue;
    if (count_[s] + v.size() > config_.queue_capacity)
      throw std::length_error("em: enqueue exceeds the device queue capacity");
    std::size_t const n = v.size();
    std::vector<std::int32_t> i32[3];
    std::vector<std::uint32_t> u32[2];
    std::vector<double> f64[9];
    std::vector<std::uint64_t> u64[3];
    for (auto& a : i32) a.resize(n);
    for (auto& a : u32) a.resize(n);
    for (auto& a : f64) a.resize(n);
    for (auto& a : u64) a.resize(n);
    for (std::size_t k = 0; k < n; ++k) {
      auto const& p = v[k];
      i32[0][k] = p.pid;
      i32[1][k] = p.medium_id;
      i32[2][k] = p.environment_node;
      u32[0][k] = p.generation;
      u32[1][k] = p.reserved;
      f64[0][k] = p.energy_gev;
      f64[1][k] = p.position_m[0];
      f64[2][k] = p.position_m[1];
      f64[3][k] = p.position_m[2];
      f64[4][k] = p.direction[0];
      f64[5][k] = p.direction[1];
      f64[6][k] = p.direction[2];
      f64[7][k] = p.time_s;
      f64[8][k] = p.weight;
      u64[0][k] = p.history_id;
      u64[1][k] = p.parent_history_id;
      u64[2][k] = p.step_id;
      double const scaled = p.energy_gev * p.weight * config_.profile_energy_scale;
      injected_ += static_cast<long double>(std::llrint(scaled));
    }
    ParticleQueue const q = queues_[s][current_[s]]->view();
    std::uint64_t const at = count_[s];
    auto put = [&](void* dst, void const* src, std::size_t elem) {
      C8MI_HIP(hipMemcpy(static_cast<unsigned char*>(dst) + at * elem, src, n * elem, hipMemcpyHostToDevice));
    };
    put(q.pid, i32[0].data(), 4);
    put(q.medium, i32[1].data(), 4);
    put(q.node, i32[2].data(), 4);
    put(q.generation, u32[0].data(), 4);
    put(q.reserved, u32[1].data(), 4);
    put(q.energy, f64[0].data(), 8);
    put(q.x, f64[1].data(), 8);
    put(q.y, f64[2].data(), 8);
    put(q.z, f64[3].data(), 8);
    put(q.dx, f64[4].data(), 8);
    put(q.dy, f64[5].data(), 8);
    put(q.dz, f64[6].data(), 8);
    put(q.time, f64[7].data(), 8);
    put(q.weight, f64[8].data(), 8);
    put(q.history, u64[0].data(), 8);
    put(q.parent, u64[1].data(), 8);
    put(q.step, u64[2].data(), 8);
    count_[s] += n;
  }
}

void EmBackend::resize_scratch(std::uint64_t n) {
  if (n <= scratch_capacity_) return;
  std::uint64_t cap = scratch_capacity_ ? scratch_capacity_ : 4096;
  while (cap < n) cap *= 2;
  staging_.allocate(2 * cap);
  packed_.allocate(cap);
  block_counts_.allocate(block_count_bytes(cap));
  offsets_.allocate(offset_bytes(cap) / sizeof(unsigned long long));
  if (!totals_.data()) totals_.allocate(1);
  scratch_capacity_ = cap;
}

void EmBackend::front(bool photon, c8mi_em_run_result& out) {
  int const s = photon ? 0 : 1, o = 1 - s;
  std::uint64_t const n = count_[s];
  if (n == 0) return;
  std::uint64_t const cap = config_.queue_capacity;
  if (2 * n > cap || count_[o] + 2 * n > cap)
    throw std::length_error("em: a front's outputs could exceed the device queue capacity");
  resize_scratch(n);
  ParticleQueue const in = queues_[s][current_[s]]->view();
  ParticleQueue const next = queues_[s][1 - current_[s]]->view();
  ParticleQueue const other = queues_[o][current_[o]]->view();
  if (timing_) C8MI_HIP(hipEventRecord(events_[0], stream_));
  launch_step(photon, context_.data(), config_.profile_bins, in, n, staging_.data(), packed_.data(),
              block_counts_.data(), stream_);
  if (timing_) C8MI_HIP(hipEventRecord(events_[1], stream_));
  launch_scan(block_counts_.data(), n, offsets_.data(), totals_.data(), stream_);
  if (timing_) C8MI_HIP(hipEventRecord(events_[2], stream_));
  launch_scatter(packed_.data(), n, staging_.data(), offsets_.data(), next, other, count_[o], next_history_,
                 stream_);
  if (timing_) C8MI_HIP(hipEventRecord(events_[3], stream_));
  C8MI_HIP(hipMemcpyAsync(host_totals_.data(), totals_.data(), sizeof(FrontTotals), hipMemcpyDeviceToHost, stream_));
  C8MI_HIP(hipStreamSynchronize(stream_));
  FrontTotals const t = *host_totals_.data();
  if (timing_) {
    float a = 0.f, b = 0.f, c = 0.f;
    C8MI_HIP(hipEventElapsedTime(&a, events_[0], events_[1]));
    C8MI_HIP(hipEventElapsedTime(&b, events_[1], events_[2]));
    C8MI_HIP(hipEventElapsedTime(&c, events_[2], events_[3]));
    ms_[photon ? 0 : 1] += a;
    launches_[photon ? 0 : 1] += 1;
    ms_[2] += b;
    launches_[2] += 1;
    ms_[3] += c;
    launches_[3] += 1;
  }
  count_[s] = t.next;
  current_[s] ^= 1;
  count_[o] += t.other;
  next_history_ += t.children;
  if (photon) {
    out.photon_fronts += 1;
    out.photon_steps += t.steps;
    out.photon_interactions += t.interactions;
  } else {
    out.lepton_fronts += 1;
    out.lepton_steps += t.steps;
    out.lepton_interactions += t.interactions;
  }
  out.secondaries += t.children;
  if (n > out.max_front) out.max_front = n;
}

void EmBackend::run(int species, std::uint64_t maximum_fronts, c8mi_em_run_result& out) {
  std::memset(&out, 0, sizeof out);
  for (std::uint64_t round = 0; round < maximum_fronts; ++round) {
    if (species == 0) {
      if (count_[0] == 0 && count_[1] == 0) break;
      if (count_[0]) front(true, out);
      if (count_[1]) front(false, out);
    } else {
      int const s = species == 1 ? 0 : 1;
      if (count_[s] == 0) break;
      front(s == 0, out);
    }
  }
  unsigned long long records[2];
  C8MI_HIP(hipMemcpy(records, record_counts_.data(), sizeof records, hipMemcpyDeviceToHost));
  out.observations = records[0];
  out.fallbacks = records[1];
  out.pending_photons = count_[0];
  out.pending_leptons = count_[1];
  out.next_history_id = next_history_;
}

void EmBackend::pending(std::uint64_t& photons, std::uint64_t& leptons) const {
  photons = count_[0];
  leptons = count_[1];
}

void EmBackend::download_profile(double* photons, double* electrons, double* positrons, double* energy_loss,
                                 c8mi_em_ledger* ledger) {
  C8MI_HIP(hipStreamSynchronize(stream_));
  std::uint32_t const bins = config_.profile_bins;
  std::vector<long long> hist(4u * (bins ? bins : 1u));
  C8MI_HIP(hipMemcpy(hist.data(), profile_.data(), hist.size() * sizeof(long long), hipMemcpyDeviceToHost));
  double* outs[4] = {photons, electrons, positrons, energy_loss};
  for (int h = 0; h < 4; ++h) {
    if (!outs[h]) continue;
    double const scale = h == 3 ? config_.profile_energy_scale : config_.profile_weight_scale;
    for (std::uint32_t k = 0; k < bins; ++k) outs[h][k] = static_cast<double>(hist[h * bins + k]) / scale;
  }
  if (!ledger) return;
  std::vector<long long> c(kCounters);
  C8MI_HIP(hipMemcpy(c.data(), counters_.data(), kCounters * sizeof(long long), hipMemcpyDeviceToHost));
  c[kInjected] += static_cast<long long>(injected_);
  sync_pending_energy_to(c.data());
  double const es = config_.profile_energy_scale;
  double* energy = &ledger->injected_energy;
  for (int k = 0; k < kEnergyTerms; ++k) energy[k] = static_cast<double>(c[k]) / es;
  ledger->steps = static_cast<std::uint64_t>(c[kSteps]);
  ledger->photon_cuts = static_cast<std::uint64_t>(c[kPhotonCuts]);
  for (int k = 0; k < 8; ++k) ledger->lepton_limits[k] = static_cast<std::uint64_t>(c[kLeptonLimit0 + k]);
  ledger->thinning_hillas_vertices = static_cast<std::uint64_t>(c[kThinHillas]);
  ledger->thinning_statistical_vertices = static_cast<std::uint64_t>(c[kThinStatistical]);
  ledger->thinning_particles_discarded = static_cast<std::uint64_t>(c[kThinDiscarded]);
  ledger->deposited_steps = static_cast<std::uint64_t>(c[kDepositedSteps]);
  ledger->fixed_point_overflows = static_cast<std::uint64_t>(c[kOverflows]);
  ledger->invalid_records = static_cast<std::uint64_t>(c[kInvalid]);
  for (int k = 0; k < 32; ++k) ledger->fallback_reasons[k] = static_cast<std::uint64_t>(c[kFallbackReason0 + k]);
}

// Weighted energy still queued (photons first, then leptons, in queue order).
void EmBackend::sync_pending_energy_to(long long* counters) {
  double pending = 0.;
  for (int s = 0; s < 2; ++s) {
    std::uint64_t const n = count_[s];
    if (n == 0) continue;
    ParticleQueue const q = queues_[s][current_[s]]->view();
    std::vector<double> e(n), w(n);
    C8MI_HIP(hipMemcpy(e.data(), q.energy, n * 8, hipMemcpyDeviceToHost));
    C8MI_HIP(hipMemcpy(w.data(), q.weight, n * 8, hipMemcpyDeviceToHost));
    for (std::uint64_t k = 0; k < n; ++k) pending += e[k] * w[k];
  }
  counters[kPending] += std::llrint(pending * config_.profile_energy_scale);
}

std::uint64_t EmBackend::download_observations(c8mi_particle_state* out, std::uint64_t capacity) {
  C8MI_HIP(hipStreamSynchronize(stream_));
  unsigned long long n = 0;
  C8MI_HIP(hipMemcpy(&n, record_counts_.data(), sizeof n, hipMemcpyDeviceToHost));
  std::uint64_t const stored = n < config_.observation_capacity ? n : config_.observation_capacity;
  std::uint64_t const copy = stored < capacity ? stored : capacity;
  if (out && copy)
    C8MI_HIP(hipMemcpy(out, observations_.data(), copy * sizeof(c8mi_particle_state), hipMemcpyDeviceToHost));
  return n;
}

std::uint64_t EmBackend::download_fallbacks(c8mi_em_fallback* out, std::uint64_t capacity) {
  C8MI_HIP(hipStreamSynchronize(stream_));
  unsigned long long n = 0;
  C8MI_HIP(hipMemcpy(&n, record_counts_.data() + 1, sizeof n, hipMemcpyDeviceToHost));
  std::uint64_t const stored = n < config_.fallback_capacity ? n : config_.fallback_capacity;
  std::uint64_t const copy = stored < capacity ? stored : capacity;
  if (out && copy)
    C8MI_HIP(hipMemcpy(out, fallbacks_.data(), copy * sizeof(c8mi_em_fallback), hipMemcpyDeviceToHost));
  return n;
}

void EmBackend::enable_timing(bool enable) {
  timing_ = enable;
  ms_.fill(0.);
  launches_.fill(0);
}

void EmBackend::kernel_time(double ms[4], std::uint64_t launches[4]) const {
  for (int k = 0; k < 4; ++k) {
    ms[k] = ms_[k];
    launches[k] = launches_[k];
  }
}

}  // namespace c8mi::em
