#include "em/em_backend.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <thread>
#include <cstring>
#include <stdexcept>
#include <string>

#include "em/em_front.hpp"
#include "em/moliere_table.hpp"

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

}  // namespace

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
    require(std::isfinite(m.magnetic_field_t[0]) && std::isfinite(m.magnetic_field_t[1]) &&
                std::isfinite(m.magnetic_field_t[2]),
            "magnetic field must be finite");
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
  require(c.shower_slots <= kMaxShowerSlots, "shower slots above the supported maximum (4096)");
  require(c.moliere_enabled <= 1, "moliere_enabled must be 0 or 1");
  require(c.moliere_fast_iterations <= 100, "moliere_fast_iterations above the Newton limit (100)");
  if (c.moliere_enabled) {
    require(valid_moliere_snapshot(c.electron_moliere), "invalid Moliere snapshot");
    require(c.electron_moliere.component_count <= 4, "Moliere snapshots of more than 4 components are not supported");
    require(c.electron_moliere_interpolation == nullptr || valid_moliere_table(*c.electron_moliere_interpolation),
            "Moliere interpolation table with non-finite entries");
  }
}

namespace {

void validate_axis(c8mi_em_shower_axis const& a) {
  require(a.grammage_g_per_cm2 != nullptr && a.support_count > 0, "shower axis projection table");
  require(std::isfinite(a.step_m) && a.step_m > 0., "shower axis projection step");
  for (int k = 0; k < 3; ++k)
    require(std::isfinite(a.start_m[k]) && std::isfinite(a.direction[k]), "shower axis start / direction");
}

std::uint64_t nonzero(std::uint64_t n) { return n ? n : 1; }

}  // namespace

void fill_context(c8mi_em_config const& config, c8mi_environment const& environment, EmContext& h) {
  h.node_count = environment.node_count;
  h.model_count = environment.model_count;
  h.observation_enabled = environment.observation_enabled;
  for (int a = 0; a < 3; ++a) {
    h.observation_point[a] = environment.observation_point_m[a];
    h.observation_normal[a] = environment.observation_normal[a];
  }
  h.maximum_magnetic_deflection = environment.maximum_magnetic_deflection_rad;
  h.seed = config.random_seed;
  h.em_cut_mev = config.em_transport_cut_mev;
  h.muon_cut_mev = config.muon_transport_cut_mev;
  h.pair_lpm = config.photon_pair_lpm;
  h.brems_lpm = config.brems_lpm;
  h.thinning = config.thinning;
  h.moliere = config.electron_moliere;
  h.moliere_enabled = config.moliere_enabled;
  h.moliere_fast_iterations = config.moliere_fast_iterations ? config.moliere_fast_iterations : 12;
  for (int i = 0; i < 70; ++i) {
    h.moliere_i1[i] = config.electron_moliere.c1[i] / (2. * i + 1.);
    h.moliere_i2[i] = config.electron_moliere.c2[i] / (2. * i + 1.);
  }
  for (int i = 0; i < 13; ++i) {
    h.moliere_i2_large[i] = i ? -0.5 * config.electron_moliere.c2_large[i] / i : 0.;
    h.moliere_half_over[i] = i ? 0.5 / i : 0.;
  }
  h.profile_bins = config.profile_bins;
  h.crossing_mode = config.profile_crossing_mode;
  h.bin_width = config.profile_bin_width_g_per_cm2;
  h.energy_loss_threshold = config.profile_energy_loss_threshold_g_per_cm2;
  h.weight_scale = config.profile_weight_scale;
  h.energy_scale = config.profile_energy_scale;
  h.observation_capacity = config.observation_capacity;
  h.fallback_capacity = config.fallback_capacity;
}

QueueStorage::QueueStorage(std::uint64_t capacity) : capacity_(capacity) {
  // 4-byte fields first, then 8-byte fields, each array 256-byte aligned
  std::size_t const n = capacity;
  auto round = [](std::size_t b) { return (b + 255u) & ~std::size_t{255u}; };
  std::size_t const b4 = round(4 * n), b8 = round(8 * n);
  storage_.allocate(6 * b4 + 12 * b8);
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
  take4(view_.shower);
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
    : config_(config), tables_(tables) {
  validate_config(config);
  validate_environment(environment);
  slots_ = config.shower_slots ? config.shower_slots : 1u;
  config_.shower_slots = slots_;
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
  std::uint64_t const bins = nonzero(config.profile_bins);
  counters_.allocate(slots_ * static_cast<std::uint64_t>(kCounters));
  profile_.allocate(slots_ * 4u * bins);
  observations_.allocate(slots_ * nonzero(config.observation_capacity));
  fallbacks_.allocate(slots_ * nonzero(config.fallback_capacity));
  record_counts_.allocate(2u * slots_);
  first_interactions_.allocate(slots_);
  first_candidates_.allocate(slots_);
  history_.allocate(slots_);
  stats_.allocate(slots_ * static_cast<std::uint64_t>(kSlotStats));
  shower_totals_.allocate(static_cast<std::size_t>(kMaxShowerSlots) * kMaxRuns * kShowerFields);
  adjust_.allocate(slots_ * static_cast<std::uint64_t>(kMaxRuns) * 3u);
  report_.allocate(kReportHeader + kReportPerSlot * static_cast<std::size_t>(slots_));
  host_report_ = PinnedBuffer<unsigned long long>(kReportHeader + kReportPerSlot * static_cast<std::size_t>(slots_));
  control_dev_.allocate(slots_);
  control_host_ = PinnedBuffer<SlotControl>(slots_);
  for (auto* b : {&history_, &stats_, &shower_totals_, &record_counts_, &adjust_, &first_candidates_})
    C8MI_HIP(hipMemset(b->data(), 0, b->bytes()));
  C8MI_HIP(hipMemset(counters_.data(), 0, counters_.bytes()));
  C8MI_HIP(hipMemset(profile_.data(), 0, profile_.bytes()));
  for (auto& species : queues_)
    for (auto& q : species) q = std::make_unique<QueueStorage>(config.queue_capacity);

  ShowerSlot base{};
  for (int a = 0; a < 3; ++a) {
    base.axis_start[a] = config.axis_start_m[a];
    base.axis_direction[a] = config.axis_direction[a];
  }
  base.axis_step = config.axis_step_m;
  base.axis_grammage = axis_.data();
  base.axis_support_count = config.axis_support_count;
  host_showers_.assign(slots_, base);
  showers_.allocate(slots_);
  C8MI_HIP(hipMemcpy(showers_.data(), host_showers_.data(), slots_ * sizeof(ShowerSlot), hipMemcpyHostToDevice));
  slot_host_ = std::vector<SlotHost>(slots_);

  EmContext& h = host_context_;
  fill_context(config, environment, h);
  h.tables = tables_.view();
  h.nodes = nodes_.data();
  h.children = children_.data();
  h.models = models_.data();
  if (config.moliere_enabled && config.electron_moliere_interpolation) {
    c8mi_moliere_interpolation const& t = *config.electron_moliere_interpolation;
    constexpr std::size_t cubics = sizeof(t.polynomials) / sizeof(t.polynomials[0]);
    constexpr std::size_t guesses = sizeof(t.initial_guess_delta) / sizeof(double);
    moliere_polynomials_.allocate(cubics);
    moliere_guess_.allocate(guesses);
    C8MI_HIP(hipMemcpy(moliere_polynomials_.data(), t.polynomials, sizeof(t.polynomials), hipMemcpyHostToDevice));
    C8MI_HIP(hipMemcpy(moliere_guess_.data(), t.initial_guess_delta, sizeof(t.initial_guess_delta),
                       hipMemcpyHostToDevice));
    h.moliere_polynomials = moliere_polynomials_.data();
    h.moliere_guess = moliere_guess_.data();
  }
  h.showers = showers_.data();
  h.shower_slots = slots_;
  h.observations = observations_.data();
  h.fallbacks = fallbacks_.data();
  h.record_counts = record_counts_.data();
  h.first_interactions = first_interactions_.data();
  h.first_candidates = first_candidates_.data();
  h.counters = counters_.data();
  h.profile = profile_.data();
  context_.allocate(1);
  C8MI_HIP(hipMemcpy(context_.data(), &h, sizeof(EmContext), hipMemcpyHostToDevice));
  vertex_count_.allocate(1);
  C8MI_HIP(hipMemset(vertex_count_.data(), 0, vertex_count_.bytes()));
  init_dev_.allocate(slots_);
  init_host_ = PinnedBuffer<SlotInit>(slots_);
  gather_slots_dev_.allocate(slots_);
  gather_slots_host_ = PinnedBuffer<std::uint32_t>(slots_);
  gather_dev_.allocate(static_cast<std::size_t>(slots_) * kGatherWords);
  gather_host_ = PinnedBuffer<unsigned long long>(static_cast<std::size_t>(slots_) * kGatherWords);
  for (auto& e : events_) C8MI_HIP(hipEventCreate(&e));
  C8MI_HIP(hipEventCreateWithFlags(&init_copied_, hipEventDisableTiming));
  C8MI_HIP(hipEventCreateWithFlags(&unstage_copied_, hipEventDisableTiming));
  C8MI_HIP(hipEventCreateWithFlags(&control_copied_, hipEventDisableTiming));
  // media a particle can carry: the environment's, and any a host particle brings
  for (std::uint32_t i = 0; i < environment.model_count; ++i) media_.push_back(environment.models[i].medium_id);
  std::sort(media_.begin(), media_.end());
  media_.erase(std::unique(media_.begin(), media_.end()), media_.end());
  order_.set_media(media_.data(), static_cast<std::uint32_t>(media_.size()), stream_);
  reset_pool();
}

void EmBackend::add_medium(std::int32_t medium) {
  auto const at = std::lower_bound(media_.begin(), media_.end(), medium);
  if (at != media_.end() && *at == medium) return;
  media_.insert(at, medium);
  order_.set_media(media_.data(), static_cast<std::uint32_t>(media_.size()), stream_);
}

PoolArrays EmBackend::pool_arrays() const {
  PoolArrays p;
  p.showers = showers_.data();
  p.profile = profile_.data();
  p.counters = counters_.data();
  p.record_counts = record_counts_.data();
  p.first_candidates = first_candidates_.data();
  p.stats = stats_.data();
  p.history = history_.data();
  p.bins = static_cast<std::uint32_t>(nonzero(config_.profile_bins));
  return p;
}

void EmBackend::start_slots(SlotInit const* inits, std::uint32_t count) {
  if (count == 0) return;
  C8MI_HIP(hipEventSynchronize(init_copied_));  // the pinned staging of the previous batch is free
  std::memcpy(init_host_.data(), inits, count * sizeof(SlotInit));
  C8MI_HIP(hipMemcpyAsync(init_dev_.data(), init_host_.data(), count * sizeof(SlotInit), hipMemcpyHostToDevice,
                          stream_));
  C8MI_HIP(hipEventRecord(init_copied_, stream_));
  launch_slot_init(pool_arrays(), init_dev_.data(), count, stream_);
  for (std::uint32_t k = 0; k < count; ++k) host_showers_[inits[k].slot] = inits[k].shower;
}

void EmBackend::activate(std::uint32_t slot, std::uint64_t shower_id, std::uint64_t first_history_id) {
  SlotHost& h = slot_host_[slot];
  h.active = true;
  h.shower_id = shower_id;
  h.fronts[0] = h.fronts[1] = 0;
  h.injected = 0;
  h.call = CallState{};
  h.call.counter = first_history_id;
  ++totals_.shower_ordinal;
}

EmBackend::~EmBackend() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (int b = 0; b < 2; ++b) {
    if (track_busy_[b]) (void)hipEventSynchronize(track_done_[b]);
    if (slot_busy_[b]) (void)hipEventSynchronize(slot_done_[b]);
    if (track_ready_[b]) (void)hipEventDestroy(track_ready_[b]);
    if (track_done_[b]) (void)hipEventDestroy(track_done_[b]);
    if (slot_done_[b]) (void)hipEventDestroy(slot_done_[b]);
  }
  slot_radio_.reset();
  for (auto& e : events_)
    if (e) (void)hipEventDestroy(e);
  if (side_stream_) (void)hipStreamSynchronize(side_stream_);
  if (open_ready_) (void)hipEventDestroy(open_ready_);
  if (open_done_) (void)hipEventDestroy(open_done_);
  if (side_stream_) (void)hipStreamDestroy(side_stream_);
  if (init_copied_) (void)hipEventDestroy(init_copied_);
  if (unstage_copied_) (void)hipEventDestroy(unstage_copied_);
  if (control_copied_) (void)hipEventDestroy(control_copied_);
  if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

void EmBackend::check_slot(std::uint32_t slot) const {
  if (slot >= slots_) throw std::invalid_argument("em: slot " + std::to_string(slot) + " outside the pool");
}

void EmBackend::update_active_range() {
  active_range_ = 0;
  for (std::uint32_t s = 0; s < slots_; ++s)
    if (slot_host_[s].active) active_range_ = s + 1;
}

void EmBackend::begin_shower(std::uint64_t shower_id, std::uint64_t first_history_id) {
  reset_pool();
  slot_begin(0, shower_id, first_history_id, nullptr);
}

// Every slot idle, every queue empty.
void EmBackend::reset_pool() {
  C8MI_HIP(hipStreamSynchronize(stream_));
  for (auto& h : slot_host_) {
    h.active = false;
    h.pending[0] = h.pending[1] = 0;
    h.staged[0] = h.staged[1] = 0;
  }
  for (int s = 0; s < 2; ++s) {
    count_[s] = 0;
    current_[s] = 0;
    clean_[s] = true;
    staged_[s].clear();
  }
  remaining_.clear();
  update_active_range();
}

void EmBackend::slot_begin(std::uint32_t slot, std::uint64_t shower_id, std::uint64_t first_history_id,
                           c8mi_em_shower_axis const* axis) {
  check_slot(slot);
  SlotHost& h = slot_host_[slot];
  if (h.pending[0] || h.pending[1] || h.staged[0] || h.staged[1])
    throw std::logic_error("em: slot " + std::to_string(slot) + " still has pending particles");
  if (slot_radio_) slot_radio_->set_slot(slot, nullptr, 0);  // no radio until the shower's observers are set
  SlotInit init{};
  ShowerSlot& s = init.shower;
  if (axis) {
    validate_axis(*axis);
    C8MI_HIP(hipStreamSynchronize(stream_));  // the slot's previous table may still be read
    upload(h.axis, axis->grammage_g_per_cm2, axis->support_count);
    for (int a = 0; a < 3; ++a) {
      s.axis_start[a] = axis->start_m[a];
      s.axis_direction[a] = axis->direction[a];
    }
    s.axis_step = axis->step_m;
    s.axis_grammage = h.axis.data();
    s.axis_support_count = axis->support_count;
  } else {
    for (int a = 0; a < 3; ++a) {
      s.axis_start[a] = config_.axis_start_m[a];
      s.axis_direction[a] = config_.axis_direction[a];
    }
    s.axis_step = config_.axis_step_m;
    s.axis_grammage = axis_.data();
    s.axis_support_count = config_.axis_support_count;
  }
  s.key = shower_id;
  init.first_history = first_history_id;
  init.slot = slot;
  start_slots(&init, 1);
  C8MI_HIP(hipStreamSynchronize(stream_));
  activate(slot, shower_id, first_history_id);
  update_active_range();
}

void EmBackend::slot_end(std::uint32_t slot) {
  check_slot(slot);
  SlotHost& h = slot_host_[slot];
  if (h.pending[0] || h.pending[1] || h.staged[0] || h.staged[1])
    throw std::logic_error("em: slot " + std::to_string(slot) + " still has pending particles");
  h.active = false;
  update_active_range();
}

bool EmBackend::can_transport(c8mi_particle_state const& p) const {
  if (!(p.pid == 22 || p.pid == 11 || p.pid == -11)) return false;
  if (!std::isfinite(p.energy_gev) || !(p.energy_gev > 0.) || !std::isfinite(p.weight) || p.weight < 0.)
    return false;
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(p.position_m[a]) || !std::isfinite(p.direction[a])) return false;
  return true;
}

void EmBackend::slot_enqueue(std::uint32_t slot, c8mi_particle_state const* particles, std::uint64_t count) {
  check_slot(slot);
  if (count == 0) return;
  if (!particles) throw std::invalid_argument("em: null particle array");
  SlotHost& h = slot_host_[slot];
  if (!h.active) throw std::logic_error("em: slot " + std::to_string(slot) + " has no shower begun");
  for (std::uint64_t i = 0; i < count; ++i)
    if (!can_transport(particles[i]))
      throw std::invalid_argument("em: particle " + std::to_string(i) + " (pid " +
                                  std::to_string(particles[i].pid) +
                                  ") cannot be transported by the device path");
  for (std::uint64_t i = 0; i < count; ++i) add_medium(particles[i].medium_id);
  totals_.particles_enqueued += count;
  for (std::uint64_t i = 0; i < count; ++i) {
    c8mi_particle_state const& p = particles[i];
    int const s = p.pid == 22 ? 0 : 1;
    Staged e{};
    e.particle = p;
    e.slot = slot;
    staged_[s].push_back(e);
    ++h.staged[s];
    h.injected += static_cast<long double>(std::llrint(p.energy_gev * p.weight * config_.profile_energy_scale));
  }
}

std::uint32_t EmBackend::active_slot_count() const {
  std::uint32_t active = 0;
  for (auto const& h : slot_host_) active += h.active ? 1u : 0u;
  return active;
}

// Uploads the host-staged particles of one species after the queue's particles
// (class host; the ordering step of the slot's next call places them after its
// pending and remaining particles, each in enqueue order): one copy from pinned
// memory and one kernel that writes the queue's fields.
void EmBackend::flush_staged(int species) {
  auto& v = staged_[species];
  if (v.empty()) return;
  std::size_t const n = v.size();
  if (count_[species] + n > config_.queue_capacity)
    throw std::length_error("em: enqueue exceeds the device queue capacity");
  C8MI_HIP(hipEventSynchronize(unstage_copied_));  // the previous batch has left the pinned buffer
  if (unstage_host_.size() < n) {
    std::size_t cap = unstage_host_.size() ? unstage_host_.size() : 1024;
    while (cap < n) cap *= 2;
    unstage_host_ = PinnedBuffer<StagedParticle>(cap);
    unstage_dev_.allocate(cap);
  }
  std::memcpy(unstage_host_.data(), v.data(), n * sizeof(StagedParticle));
  C8MI_HIP(hipMemcpyAsync(unstage_dev_.data(), unstage_host_.data(), n * sizeof(StagedParticle),
                          hipMemcpyHostToDevice, stream_));
  C8MI_HIP(hipEventRecord(unstage_copied_, stream_));
  launch_unstage(unstage_dev_.data(), n, queues_[species][current_[species]]->view(), count_[species], stream_);
  for (auto const& e : v) {
    --slot_host_[e.slot].staged[species];
    ++slot_host_[e.slot].pending[species];
  }
  count_[species] += n;
  clean_[species] = false;
  totals_.host_to_device_bytes += n * sizeof(StagedParticle);
  v.clear();
}

void EmBackend::resize_scratch(std::uint64_t n) {
  if (n <= scratch_capacity_) return;
  std::uint64_t cap = scratch_capacity_ ? scratch_capacity_ : 4096;
  while (cap < n) cap *= 2;
  staging_.allocate(2 * cap);
  packed_.allocate(cap);
  block_counts_.allocate(block_count_bytes(cap));
  offsets_.allocate(offset_bytes(cap) / sizeof(unsigned long long));
  vertices_.allocate(cap);
  if (config_.moliere_enabled) {
    moves_.allocate(cap * lepton_move_bytes());
    moved_.allocate(cap);
    moliere_open_.allocate(cap * moliere_open_bytes());
    open_vertices_.allocate(cap);
    if (!moliere_open_count_.data()) {
      moliere_open_count_.allocate(1);
      open_vertex_count_.allocate(1);
      C8MI_HIP(hipMemset(open_vertex_count_.data(), 0, sizeof(unsigned long long)));
      C8MI_HIP(hipStreamCreate(&side_stream_));
      C8MI_HIP(hipEventCreateWithFlags(&open_ready_, hipEventDisableTiming));
      C8MI_HIP(hipEventCreateWithFlags(&open_done_, hipEventDisableTiming));
    }
  }
  scratch_capacity_ = cap;
}

void EmBackend::wait_radio_tracks() {
  for (int b = 0; b < 2; ++b) {
    if (track_busy_[b]) {
      C8MI_HIP(hipEventSynchronize(track_done_[b]));
      track_busy_[b] = false;
    }
    if (slot_busy_[b]) {
      C8MI_HIP(hipEventSynchronize(slot_done_[b]));
      slot_busy_[b] = false;
    }
  }
}

void EmBackend::create_track_events() {
  if (track_ready_[0]) return;
  for (int b = 0; b < 2; ++b) {
    C8MI_HIP(hipEventCreateWithFlags(&track_ready_[b], hipEventDisableTiming));
    C8MI_HIP(hipEventCreateWithFlags(&track_done_[b], hipEventDisableTiming));
    C8MI_HIP(hipEventCreateWithFlags(&slot_done_[b], hipEventDisableTiming));
  }
}

void EmBackend::attach_radio(radio::RadioAccumulator* radio) {
  wait_radio_tracks();
  if (radio) create_track_events();
  radio_ = radio;
}

void EmBackend::enable_slot_radio(c8mi_radio_config const& config, std::uint32_t observers_per_slot,
                                  std::uint32_t bins_per_observer) {
  wait_radio_tracks();
  slot_radio_.reset();
  slot_radio_ = std::make_unique<radio::SlotRadio>(config, slots_, observers_per_slot, bins_per_observer, nullptr);
  create_track_events();
  std::size_t const keys = static_cast<std::size_t>(slots_) * kMaxRuns;
  for (int b = 0; b < 2; ++b) {
    segment_keys_[b].allocate(2 * keys);
    C8MI_HIP(hipMemset(segment_keys_[b].data(), 0, segment_keys_[b].bytes()));
    segments_[b].allocate(keys);
    segment_count_[b].allocate(1);
    item_count_[b].allocate(1);
  }
}

void EmBackend::slot_set_radio(std::uint32_t slot, c8mi_radio_observer const* observers, std::uint64_t count) {
  check_slot(slot);
  if (!slot_radio_) throw std::logic_error("em: per-shower radio is not enabled");
  slot_radio_->set_slot(slot, observers, count);
}

void EmBackend::slot_download_radio(std::uint32_t slot, std::int64_t* coreas, std::int64_t* zhs,
                                    c8mi_radio_counters* counters) {
  check_slot(slot);
  if (!slot_radio_) throw std::logic_error("em: per-shower radio is not enabled");
  slot_radio_->download(slot, coreas, zhs, counters);
}

void EmBackend::slot_radio_time(double& ms, std::uint64_t& launches) {
  ms = 0.;
  launches = 0;
  if (slot_radio_) slot_radio_->kernel_time_ms(ms, launches);
}

namespace {

constexpr std::uint64_t kPhotonCallFronts = 16;   // EmRouter.inl:368
constexpr std::uint64_t kLeptonCallFronts = 1024; // EmRouter.inl:405

std::uint32_t bits_for(std::uint64_t values) {
  std::uint32_t b = 0;
  while (b < 64 && (std::uint64_t{1} << b) < values) ++b;
  return b;
}

}  // namespace

// History reservation of a device-resident call of n particles
// (reservePhotonSecondaryRange, EmRouter.inl:640-656: 2 n per front of the
// call; reserveResidentSecondaryRange, EmRouter.inl:658-679:
// max(3 n, min(2^20, 64 * 3 n)) with its exclusive limit).
void EmBackend::reserve_call(CallState& c, int species, std::uint64_t n, c8mi_em_run_result& out) const {
  std::uint64_t count;
  if (species == 0) {
    if (n > std::numeric_limits<std::uint64_t>::max() / (2 * kPhotonCallFronts))
      throw std::overflow_error("em: photon history reservation overflow");
    count = n * 2 * kPhotonCallFronts;
  } else {
    if (n > std::numeric_limits<std::uint64_t>::max() / 3)
      throw std::overflow_error("em: lepton history reservation overflow");
    std::uint64_t const minimum = 3 * n;
    count = minimum;
    if (minimum <= std::numeric_limits<std::uint64_t>::max() / 64)
      count = std::max(std::min<std::uint64_t>(std::uint64_t{1} << 20, minimum * 64), minimum);
  }
  if (count > std::numeric_limits<std::uint64_t>::max() - c.counter)
    throw std::overflow_error("em: history range limit overflow");
  c.next = c.counter;
  c.limit = c.counter + count;
  c.counter += count;
  c.open = c.opening = true;
  c.call_fronts = 0;
  c.cap = species == 0 ? kPhotonCallFronts : kLeptonCallFronts;
  c.min_batch = config_.resident_minimum_batch;
  out.reserved_history_ids += count;
  ++(species == 0 ? out.photon_calls : out.lepton_calls);
}

// The loop head of run_photon_cascade / run_lepton_cascade for one shower
// (PhotonCascade.inl:286-300, LeptonCascade.inl:240-252), applied until the
// shower's next front is known: a call runs while its queue holds particles and
// fewer than `cap` fronts ran; after its first front it ends below the minimum
// batch; a lepton call ends when the front could exhaust its history range. An
// ended call passes to the other species, where the device-resident schedule
// opens the next call (reserving its range) if that species has particles.
// n[s]: the shower's particles of each species. Returns true when the shower
// runs a front of species c.phase now.
bool EmBackend::settle(CallState& c, std::uint64_t const n[2], c8mi_em_run_result& out) const {
  for (int step = 0; step < 4; ++step) {
    int const s = c.phase;
    if (!c.open) {
      if (!c.resident) return false;  // a wavefront call ended: the caller decides
      if (n[s] == 0) {
        if (n[1 - s] == 0) return false;
        c.phase ^= 1;
        continue;
      }
      reserve_call(c, s, n[s], out);
    }
    int stop = -1;
    if (n[s] == 0) stop = kStopEmpty;
    else if (c.call_fronts >= c.cap) stop = kStopCap;
    else if (c.call_fronts != 0 && n[s] < c.min_batch) stop = kStopMinimumBatch;
    else if (s == 1 && n[s] > (c.limit - c.next) / 3) stop = kStopHistory;
    if (stop < 0) return true;
    if (stop == kStopMinimumBatch) ++out.below_minimum_batch_checkpoints;
    if (stop == kStopHistory) ++out.history_range_checkpoints;
    c.open = c.opening = false;
    c.stop = stop;
    c.phase ^= 1;
  }
  return false;
}

bool EmBackend::front(bool photon, c8mi_em_run_result& out) {
  int const s = photon ? 0 : 1, o = 1 - s;
  // host particles join the queue now; the ordering step of their shower's
  // call places them after its pending and remaining particles
  flush_staged(s);
  std::uint64_t const n = count_[s];
  // Decide every active shower's part in this front on copies of their call
  // states; they are committed once nothing can refuse the front.
  plan_.assign(active_range_, CallState{});
  front_counts_.assign(active_range_, 0);
  c8mi_em_run_result decided{};
  std::uint64_t m = 0;
  bool any_open = false, any_bucket = false;
  C8MI_HIP(hipEventSynchronize(control_copied_));  // the previous front's control upload has left
  for (std::uint32_t k = 0; k < active_range_; ++k) {
    SlotControl& c = control_host_.data()[k];
    c = SlotControl{};
    SlotHost const& h = slot_host_[k];
    plan_[k] = h.call;
    if (!h.active) continue;
    std::uint64_t const counts[2] = {h.pending[0] + h.staged[0], h.pending[1] + h.staged[1]};
    if (!settle(plan_[k], counts, decided) || plan_[k].phase != s) continue;
    std::uint64_t const cnt = h.pending[s];
    c.flags = kControlRun;
    if (plan_[k].opening) {
      c.flags |= kControlOpen;
      c.first_history = plan_[k].next;
      any_open = true;
    }
    if (cnt >= kBucketMinimum) {
      c.flags |= kControlBucket;
      any_bucket = true;
    }
    front_counts_[k] = cnt;
    m += cnt;
  }
  auto commit = [&] {
    for (std::uint32_t k = 0; k < active_range_; ++k) {
      if (!slot_host_[k].active) continue;
      slot_host_[k].call = plan_[k];
      if (control_host_.data()[k].flags & kControlRun) {
        slot_host_[k].call.opening = false;
        ++slot_host_[k].call.call_fronts;
      }
    }
    out.photon_calls += decided.photon_calls;
    out.lepton_calls += decided.lepton_calls;
    out.reserved_history_ids += decided.reserved_history_ids;
    out.below_minimum_batch_checkpoints += decided.below_minimum_batch_checkpoints;
    out.history_range_checkpoints += decided.history_range_checkpoints;
    totals_.photon_calls += decided.photon_calls;
    totals_.lepton_calls += decided.lepton_calls;
    totals_.reserved_history_ids += decided.reserved_history_ids;
    totals_.below_minimum_batch_checkpoints += decided.below_minimum_batch_checkpoints;
    totals_.history_range_checkpoints += decided.history_range_checkpoints;
  };
  if (m == 0) {
    commit();
    return false;
  }
  bool const parked = m < n;
  bool const ordered = any_open || any_bucket || parked || !clean_[s];
  std::uint64_t const cap = config_.queue_capacity;
  // Everything that can refuse this front is checked before the first launch:
  // once the kernels run, the device ledger, profile and queues move on.
  if (n + m > cap || count_[o] + 2 * m > cap)
    throw std::length_error("em: a front's outputs could exceed the device queue capacity");
  if (!photon) {
    if (radio_) radio_->check_accumulate(m);
    if (slot_radio_) slot_radio_->check_project(m, slot_radio_->chunk_for(m), active_range_ * kMaxRuns);
  }
  resize_scratch(m);
  commit();
  ParticleQueue const cur = queues_[s][current_[s]]->view();
  ParticleQueue const alt = queues_[s][1 - current_[s]]->view();
  if (ordered) {
    C8MI_HIP(hipMemcpyAsync(control_dev_.data(), control_host_.data(), active_range_ * sizeof(SlotControl),
                            hipMemcpyHostToDevice, stream_));
    C8MI_HIP(hipEventRecord(control_copied_, stream_));
    totals_.host_to_device_bytes += active_range_ * sizeof(SlotControl);
    OrderLayout layout{};
    layout.class_bits = any_open ? 2u : 0u;
    layout.medium_bits = order_.medium_bits();
    layout.bucket_bits = any_bucket ? 18u + layout.medium_bits : 0u;
    layout.slot_bits = bits_for(active_range_);
    layout.parked = parked ? 1u : 0u;
    order_.order(cur, alt, n, m, control_dev_.data(), active_range_, history_.data(), layout, stream_);
    ++totals_.ordered_fronts;
    if (parked) {
      ++totals_.parked_fronts;
      totals_.parked_particles += n - m;
    }
  }
  // ordered: the front reads the gathered buffer and writes its continuing
  // particles back into the current one; else it reads the current buffer
  ParticleQueue const front_in = ordered ? alt : cur;
  ParticleQueue const next = ordered ? cur : alt;
  ParticleQueue const other = queues_[o][current_[o]]->view();
  for (std::uint32_t k = 0; k < active_range_; ++k) {
    if (!(control_host_.data()[k].flags & kControlRun)) continue;
    ++slot_host_[k].fronts[s];
    bool const bucketed = front_counts_[k] >= kBucketMinimum;
    out.bucketed_fronts += bucketed ? 1u : 0u;
    out.bucketed_particles += bucketed ? front_counts_[k] : 0u;
    (bucketed ? totals_.bucketed_fronts : totals_.small_fronts) += 1;
    (bucketed ? totals_.bucketed_particles : totals_.small_particles) += front_counts_[k];
  }
  FrontArgs f{};
  f.in = front_in;
  f.count = m;
  f.runs = 1;
  f.run_end[0] = m;
  f.slots = active_range_;
  f.staging = staging_.data();
  f.packed = packed_.data();
  f.block_counts = block_counts_.data();
  f.offsets = offsets_.data();
  f.shower_totals = shower_totals_.data();
  f.adjust = adjust_.data();
  f.history = history_.data();
  f.stats = stats_.data();
  f.report = report_.data();
  f.vertices = vertices_.data();
  f.vertex_count = vertex_count_.data();
  f.moves = moves_.data();
  f.moved = moved_.data();
  f.moliere_open = moliere_open_.data();
  f.moliere_open_count = moliere_open_count_.data();
  f.open_vertices = open_vertices_.data();
  f.open_vertex_count = open_vertex_count_.data();
  f.first_interactions = first_interactions_.data();
  f.first_candidates = first_candidates_.data();
  int const track_buffer = track_parity_;
  if ((radio_ || slot_radio_) && !photon) {
    // tracks [9][n]; the buffer's previous projections must be done first
    std::uint64_t& capacity = track_capacity_[track_buffer];
    if (capacity < m) {
      if (track_busy_[track_buffer]) C8MI_HIP(hipEventSynchronize(track_done_[track_buffer]));
      if (slot_busy_[track_buffer]) C8MI_HIP(hipEventSynchronize(slot_done_[track_buffer]));
      track_busy_[track_buffer] = slot_busy_[track_buffer] = false;
      capacity = std::max<std::uint64_t>(m, 2 * capacity);
      tracks_[track_buffer].allocate(C8MI_RADIO_TRACK_FIELDS * capacity);
    }
    if (track_busy_[track_buffer]) C8MI_HIP(hipStreamWaitEvent(stream_, track_done_[track_buffer], 0));
    if (slot_busy_[track_buffer]) C8MI_HIP(hipStreamWaitEvent(stream_, slot_done_[track_buffer], 0));
    f.tracks = tracks_[track_buffer].data();
    f.track_stride = m;
  }
  // Projects this front's tracks on the accumulators' streams once the step kernel is done.
  auto project_tracks = [&] {
    if (!f.tracks) return;
    std::uint64_t const chunk = slot_radio_ ? slot_radio_->chunk_for(m) : 0;
    if (slot_radio_) {
      std::size_t const keys = static_cast<std::size_t>(slots_) * kMaxRuns;
      unsigned long long* key_begin = segment_keys_[track_buffer].data();
      launch_radio_segments(f, active_range_, key_begin, key_begin + keys, slot_radio_->observer_counts(), chunk,
                            slot_radio_->bin_tiles(), segments_[track_buffer].data(),
                            segment_count_[track_buffer].data(), item_count_[track_buffer].data(), stream_);
    }
    C8MI_HIP(hipEventRecord(track_ready_[track_buffer], stream_));
    if (radio_) {
      C8MI_HIP(hipStreamWaitEvent(radio_->stream(), track_ready_[track_buffer], 0));
      radio_->accumulate_tracks(f.tracks, m, m);
      C8MI_HIP(hipEventRecord(track_done_[track_buffer], radio_->stream()));
      track_busy_[track_buffer] = true;
    }
    if (slot_radio_) {
      C8MI_HIP(hipStreamWaitEvent(slot_radio_->stream(), track_ready_[track_buffer], 0));
      slot_radio_->project(f.tracks, m, m, chunk, segments_[track_buffer].data(),
                           segment_count_[track_buffer].data(), item_count_[track_buffer].data(),
                           active_range_ * kMaxRuns);
      C8MI_HIP(hipEventRecord(slot_done_[track_buffer], slot_radio_->stream()));
      slot_busy_[track_buffer] = true;
    }
    track_parity_ ^= 1;
  };
  // the launches of this front, each followed by an event when timing
  int kinds[kFrontKernels + 4];
  int launched = 0;
  auto mark = [&](int kind) {
    if (!timing_) return;
    kinds[launched] = kind;
    C8MI_HIP(hipEventRecord(events_[++launched], stream_));
  };
  if (timing_) C8MI_HIP(hipEventRecord(events_[0], stream_));
  if (photon) {
    launch_photon_step(context_.data(), config_.profile_bins, f, stream_);
    mark(kKernPhotonStep);
  } else if (config_.moliere_enabled) {
    // Molière scattering runs in the finish kernel, with its own register budget
    launch_lepton_move(context_.data(), config_.profile_bins, f, stream_);
    mark(kKernLeptonTransport);
    bool const table = host_context_.moliere_polynomials != nullptr;
    launch_lepton_moliere(context_.data(), table, f, stream_);
    mark(kKernLeptonTransport);
    // the few open samples continue on the side stream while the other
    // sources finish and run their vertices here
    C8MI_HIP(hipEventRecord(open_ready_, stream_));
    C8MI_HIP(hipStreamWaitEvent(side_stream_, open_ready_, 0));
    launch_lepton_open(context_.data(), config_.profile_bins, table, f, side_stream_);
    C8MI_HIP(hipEventRecord(open_done_, side_stream_));
    launch_lepton_finish(context_.data(), config_.profile_bins, f, stream_);
    mark(kKernLeptonTransport);
    launch_lepton_vertex(context_.data(), config_.profile_bins, f, stream_);
    mark(kKernLeptonVertex);
    C8MI_HIP(hipStreamWaitEvent(stream_, open_done_, 0));
    mark(kKernLeptonTransport);  // what the open samples add beyond the overlap
    project_tracks();
  } else {
    launch_lepton_transport(context_.data(), config_.profile_bins, f, stream_);
    mark(kKernLeptonTransport);
    project_tracks();
    launch_lepton_vertex(context_.data(), config_.profile_bins, f, stream_);
    mark(kKernLeptonVertex);
  }
  launch_count(f, stream_);
  mark(kKernCount);
  launch_scan(photon, f, stream_);
  mark(kKernScan);
  launch_scatter(f, next, other, count_[o], stream_);
  mark(kKernScatter);
  // showers that sat this front out follow the continuing particles unchanged
  if (parked) launch_carry(front_in, m, n, next, report_.data(), stream_);
  C8MI_HIP(hipMemcpyAsync(host_report_.data(), report_.data(),
                          (kReportHeader + kReportPerSlot * static_cast<std::size_t>(active_range_)) *
                              sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, stream_));
  C8MI_HIP(hipStreamSynchronize(stream_));
  FrontTotals t;
  std::memcpy(&t, host_report_.data(), sizeof t);
  totals_.device_to_host_bytes += (kReportHeader + kReportPerSlot * active_range_) * sizeof(unsigned long long);
  for (int k = 0; k < launched; ++k) {
    float ms = 0.f;
    C8MI_HIP(hipEventElapsedTime(&ms, events_[k], events_[k + 1]));
    ms_[kinds[k]] += ms;
    launches_[kinds[k]] += 1;
  }
  unsigned long long const* per_slot = host_report_.data() + kReportHeader;
  for (std::uint32_t k = 0; k < active_range_; ++k) {
    if (!(control_host_.data()[k].flags & kControlRun)) continue;
    SlotHost& h = slot_host_[k];
    h.pending[s] = per_slot[kReportPerSlot * k];
    h.pending[o] += per_slot[kReportPerSlot * k + 1];
    h.call.next += per_slot[kReportPerSlot * k + 2];
  }
  count_[s] = t.next + (n - m);
  if (!ordered) current_[s] ^= 1;
  clean_[s] = !parked;
  count_[o] += t.other;
  if (t.other) clean_[o] = false;
  if (photon) {
    out.photon_fronts += 1;
    out.photon_steps += t.steps;
    out.photon_interactions += t.interactions;
    totals_.photon_wavefronts += 1;
  } else {
    out.lepton_fronts += 1;
    out.lepton_steps += t.steps;
    out.lepton_interactions += t.interactions;
    totals_.lepton_wavefronts += 1;
  }
  out.secondaries += t.children;
  if (m > out.max_front) out.max_front = m;
  totals_.wavefronts += 1;
  totals_.particles_advanced += t.steps;
  totals_.particles_produced += t.children;
  totals_.interactions_selected += t.interactions;
  return true;
}

void EmBackend::run(std::uint64_t maximum_rounds, c8mi_em_run_result& out) {
  std::memset(&out, 0, sizeof out);
  run_rounds_loop(maximum_rounds, out);
  std::vector<unsigned long long> records(2u * slots_);
  C8MI_HIP(hipMemcpy(records.data(), record_counts_.data(), records.size() * sizeof(unsigned long long),
                     hipMemcpyDeviceToHost));
  for (std::uint32_t k = 0; k < active_range_; ++k) {
    out.observations += records[2 * k];
    out.fallbacks += records[2 * k + 1];
  }
  pending(out.pending_photons, out.pending_leptons);
  out.next_history_id = slot_host_[0].call.counter;
}

// Rounds of the schedule: every shower's photon front (if its call runs one
// now), then its lepton front. No record or history read-back.
void EmBackend::run_rounds_loop(std::uint64_t maximum_rounds, c8mi_em_run_result& out) {
  // Showers are independent: any interleaving that keeps each shower's own
  // front sequence gives the same results. Every round offers each shower its
  // next front (a photon front, then a lepton front), so every shower advances
  // every round; letting one species wait for a larger front instead measured
  // 1.9x more fronts on a pool of 100 GeV showers.
  for (std::uint64_t round = 0; round < maximum_rounds; ++round) {
    bool const photons = front(true, out);
    bool const leptons = front(false, out);
    if (photons || leptons) continue;
    if (count_[0] + staged_[0].size() + count_[1] + staged_[1].size() == 0) break;
    // a shower with particles always has a next front once its calls settle
    throw std::logic_error("em: particles queued but no shower's schedule runs a front");
  }
}

void EmBackend::run_call(bool photon, c8mi_particle_state const* particles, std::uint64_t count,
                         std::uint64_t first_secondary_history_id, std::uint64_t maximum_wavefronts,
                         std::uint64_t history_limit_exclusive, std::uint64_t minimum_resident_batch,
                         c8mi_em_wavefront_result& out) {
  std::memset(&out, 0, sizeof out);
  int const s = photon ? 0 : 1, o = 1 - s;
  SlotHost& h = slot_host_[0];
  if (!h.active || active_slot_count() != 1)
    throw std::logic_error("em: the wavefront calls act on the one shower begun with begin_shower");
  if (count && !particles) throw std::invalid_argument("em: null particle array");
  remaining_.clear();
  std::uint64_t const total = h.pending[s] + h.staged[s] + count;
  out.input_particles = total;
  out.other_species_pending = h.pending[o] + h.staged[o];
  if (total == 0) {  // PhotonCascade.inl:206-209
    out.completed = 1;
    return;
  }
  // PhotonCascade.inl:211-214, LeptonCascade.inl:127-132
  if (first_secondary_history_id == 0 || maximum_wavefronts == 0 || minimum_resident_batch == 0)
    throw std::invalid_argument("em: a wavefront call requires nonzero history ID, wavefronts and batch");
  if (!photon && history_limit_exclusive <= first_secondary_history_id)
    throw std::invalid_argument("em: the lepton history limit must exceed the first history ID");
  if (2 * total > config_.queue_capacity)
    throw std::length_error("em: wavefront input exceeds the advertised resident batch limit");
  slot_enqueue(0, particles, count);
  unsigned long long records_before[2] = {0, 0};
  C8MI_HIP(hipStreamSynchronize(stream_));
  C8MI_HIP(hipMemcpy(records_before, record_counts_.data(), sizeof records_before, hipMemcpyDeviceToHost));
  c8mi_em_statistics const before = totals_;
  CallState& c = h.call;
  CallState const saved = c;
  c.phase = s;
  c.open = c.opening = true;
  c.resident = false;
  c.stop = kStopEmpty;
  c.call_fronts = 0;
  c.next = first_secondary_history_id;
  c.limit = photon ? std::numeric_limits<std::uint64_t>::max() : history_limit_exclusive;
  c.cap = maximum_wavefronts;
  c.min_batch = minimum_resident_batch;
  c8mi_em_run_result acc{};
  try {
    while (c.open && c.phase == s) front(photon, acc);
  } catch (...) {
    if (c.open && c.call_fronts == 0) c = saved;  // refused before its first front
    c.resident = true;
    throw;
  }
  c.resident = true;
  out.wavefronts = c.call_fronts;
  out.transport_records = photon ? acc.photon_steps : acc.lepton_steps;
  out.interactions = photon ? acc.photon_interactions : acc.lepton_interactions;
  out.secondaries = acc.secondaries;
  out.next_secondary_history_id = c.next;
  out.peak_resident = acc.max_front;
  out.bucketed_fronts = totals_.bucketed_fronts - before.bucketed_fronts;
  out.bucketed_particles = totals_.bucketed_particles - before.bucketed_particles;
  out.small_fronts = totals_.small_fronts - before.small_fronts;
  out.small_particles = totals_.small_particles - before.small_particles;
  out.below_minimum_batch_checkpoint = c.stop == kStopMinimumBatch ? 1u : 0u;
  out.history_range_exhausted = c.stop == kStopHistory ? 1u : 0u;
  // the species' particles left on the device are the call's remaining ones
  std::uint64_t const left = count_[s];
  if (left) {
    resize_scratch(left);
    launch_pack(queues_[s][current_[s]]->view(), left, staging_.data(), stream_);
    remaining_.resize(left);
    C8MI_HIP(hipMemcpyAsync(remaining_.data(), staging_.data(), left * sizeof(c8mi_particle_state),
                            hipMemcpyDeviceToHost, stream_));
    totals_.device_to_host_bytes += left * sizeof(c8mi_particle_state);
    totals_.remaining_downloaded += left;
  }
  unsigned long long records_after[2] = {0, 0};
  C8MI_HIP(hipMemcpyAsync(records_after, record_counts_.data(), sizeof records_after, hipMemcpyDeviceToHost,
                          stream_));
  C8MI_HIP(hipStreamSynchronize(stream_));
  count_[s] = 0;
  h.pending[s] = 0;
  clean_[s] = true;
  // the remaining particles leave the device ledger: their energy no longer
  // counts as injected (it is injected again when the router sends them back)
  for (auto const& p : remaining_)
    h.injected -= static_cast<long double>(std::llrint(p.energy_gev * p.weight * config_.profile_energy_scale));
  out.remaining = left;
  out.completed = left == 0 ? 1u : 0u;
  out.other_species_pending = h.pending[o] + h.staged[o];
  out.observations = records_after[0] - records_before[0];
  out.fallbacks = records_after[1] - records_before[1];
}

void EmBackend::limits(c8mi_em_limits& out) const {
  std::memset(&out, 0, sizeof out);
  out.minimum_batch_size = config_.minimum_batch_size ? config_.minimum_batch_size : 16384;
  out.maximum_resident_photon_batch = config_.queue_capacity / 2;
  out.maximum_resident_lepton_batch = config_.queue_capacity / 2;
  out.maximum_resident_input_batch = config_.queue_capacity / 2;
  out.queue_capacity = config_.queue_capacity;
  out.shower_slots = slots_;
}

void EmBackend::statistics(c8mi_em_statistics& out) const { out = totals_; }

void EmBackend::pending(std::uint64_t& photons, std::uint64_t& leptons) const {
  photons = count_[0] + staged_[0].size();
  leptons = count_[1] + staged_[1].size();
}

void EmBackend::slot_state(std::uint32_t slot, c8mi_em_slot_state& out) const {
  check_slot(slot);
  std::memset(&out, 0, sizeof out);
  SlotHost const& h = slot_host_[slot];
  unsigned long long st[kSlotStats], records[2];
  C8MI_HIP(hipStreamSynchronize(stream_));
  C8MI_HIP(hipMemcpy(st, stats_.data() + static_cast<std::uint64_t>(slot) * kSlotStats, sizeof st,
                     hipMemcpyDeviceToHost));
  C8MI_HIP(hipMemcpy(records, record_counts_.data() + 2u * slot, sizeof records, hipMemcpyDeviceToHost));
  out.shower_id = h.shower_id;
  out.pending_photons = h.pending[0] + h.staged[0];
  out.pending_leptons = h.pending[1] + h.staged[1];
  out.next_history_id = h.call.counter;
  out.photon_fronts = h.fronts[0];
  out.lepton_fronts = h.fronts[1];
  out.photon_steps = st[0];
  out.lepton_steps = st[1];
  out.photon_interactions = st[2];
  out.lepton_interactions = st[3];
  out.secondaries = st[4];
  out.observations = records[0];
  out.fallbacks = records[1];
  out.active = h.active ? 1u : 0u;
}

// Weighted energy of a slot still queued: device queues (photons first, then
// leptons, in queue order), then host-staged particles.
long long EmBackend::pending_energy(std::uint32_t slot) {
  double pending = 0.;
  for (int s = 0; s < 2; ++s) {
    std::uint64_t const n = count_[s];
    if (n == 0 || slot_host_[slot].pending[s] == 0) continue;
    ParticleQueue const q = queues_[s][current_[s]]->view();
    std::vector<double> e(n), w(n);
    std::vector<std::uint32_t> owner(n);
    C8MI_HIP(hipMemcpy(e.data(), q.energy, n * 8, hipMemcpyDeviceToHost));
    C8MI_HIP(hipMemcpy(w.data(), q.weight, n * 8, hipMemcpyDeviceToHost));
    C8MI_HIP(hipMemcpy(owner.data(), q.shower, n * 4, hipMemcpyDeviceToHost));
    for (std::uint64_t k = 0; k < n; ++k)
      if ((owner[k] & kSlotBits) == slot) pending += e[k] * w[k];
  }
  for (int s = 0; s < 2; ++s)
    for (auto const& p : staged_[s])
      if (p.slot == slot) pending += p.particle.energy_gev * p.particle.weight;
  return std::llrint(pending * config_.profile_energy_scale);
}

void EmBackend::slot_download_profile(std::uint32_t slot, double* photons, double* electrons, double* positrons,
                                      double* energy_loss, c8mi_em_ledger* ledger) {
  check_slot(slot);
  C8MI_HIP(hipStreamSynchronize(stream_));
  std::uint32_t const bins = config_.profile_bins;
  std::uint64_t const stride = 4u * nonzero(bins);
  std::vector<long long> hist(stride);
  C8MI_HIP(hipMemcpy(hist.data(), profile_.data() + slot * stride, stride * sizeof(long long),
                     hipMemcpyDeviceToHost));
  double* outs[4] = {photons, electrons, positrons, energy_loss};
  for (int h = 0; h < 4; ++h) {
    if (!outs[h]) continue;
    double const scale = h == 3 ? config_.profile_energy_scale : config_.profile_weight_scale;
    for (std::uint32_t k = 0; k < bins; ++k) outs[h][k] = static_cast<double>(hist[h * bins + k]) / scale;
  }
  if (!ledger) return;
  std::vector<long long> c(kCounters);
  C8MI_HIP(hipMemcpy(c.data(), counters_.data() + static_cast<std::uint64_t>(slot) * kCounters,
                     kCounters * sizeof(long long), hipMemcpyDeviceToHost));
  c[kPending] += pending_energy(slot);
  decode_ledger(c.data(), slot, *ledger);
}

// Ledger of a slot from its device counters (the host adds the injected energy
// it accumulated at enqueue).
void EmBackend::decode_ledger(long long* c, std::uint32_t slot, c8mi_em_ledger& out) const {
  c[kInjected] += static_cast<long long>(slot_host_[slot].injected);
  ledger_from_counters(c, config_.profile_energy_scale, out);
}

void ledger_from_counters(long long const* c, double energy_scale, c8mi_em_ledger& out) {
  c8mi_em_ledger* ledger = &out;
  double* energy = &ledger->injected_energy;
  for (int k = 0; k < kEnergyTerms; ++k) energy[k] = static_cast<double>(c[k]) / energy_scale;
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
  ledger->native_muon_pair_cut_energy = static_cast<double>(c[kNativeMuonCut]) / energy_scale;
}

std::uint64_t EmBackend::slot_download_records(std::uint32_t slot, int kind, void* out, std::size_t record_bytes,
                                               std::uint64_t capacity) {
  check_slot(slot);
  C8MI_HIP(hipStreamSynchronize(stream_));
  unsigned long long n = 0;
  C8MI_HIP(hipMemcpy(&n, record_counts_.data() + 2u * slot + kind, sizeof n, hipMemcpyDeviceToHost));
  std::uint64_t const region = kind == 0 ? config_.observation_capacity : config_.fallback_capacity;
  std::uint64_t const stored = n < region ? n : region;
  std::uint64_t const copy = stored < capacity ? stored : capacity;
  if (out && copy) {
    unsigned char const* base = kind == 0 ? reinterpret_cast<unsigned char const*>(observations_.data())
                                          : reinterpret_cast<unsigned char const*>(fallbacks_.data());
    C8MI_HIP(hipMemcpy(out, base + slot * region * record_bytes, copy * record_bytes, hipMemcpyDeviceToHost));
  }
  return n;
}

bool EmBackend::slot_download_first_interaction(std::uint32_t slot, c8mi_em_first_interaction& out) {
  check_slot(slot);
  C8MI_HIP(hipStreamSynchronize(stream_));
  unsigned long long candidates = 0;
  C8MI_HIP(hipMemcpy(&candidates, first_candidates_.data() + slot, sizeof candidates, hipMemcpyDeviceToHost));
  if (candidates == 0) return false;
  C8MI_HIP(hipMemcpy(&out, first_interactions_.data() + slot, sizeof out, hipMemcpyDeviceToHost));
  return true;
}

void EmBackend::slot_pending(std::uint32_t slot, std::uint64_t& photons, std::uint64_t& leptons) const {
  check_slot(slot);
  SlotHost const& h = slot_host_[slot];
  photons = h.pending[0] + h.staged[0];
  leptons = h.pending[1] + h.staged[1];
}

void EmBackend::serve(c8mi_em_request const* requests, std::uint64_t count, c8mi_em_serve_options const& options,
                      c8mi_em_request_result* results, c8mi_em_run_result& totals) {
  std::memset(&totals, 0, sizeof totals);
  if (count == 0) return;
  if (!requests || !results) throw std::invalid_argument("em: null request or result array");
  for (std::uint32_t slot = 0; slot < slots_; ++slot) {
    SlotHost const& h = slot_host_[slot];
    if (h.pending[0] || h.pending[1] || h.staged[0] || h.staged[1])
      throw std::logic_error("em: serve needs a pool without pending particles");
  }
  std::uint64_t axis_words = 0;
  for (std::uint64_t k = 0; k < count; ++k) {
    c8mi_em_request const& r = requests[k];
    if (!std::isfinite(r.arrival_s) || (k && r.arrival_s < requests[k - 1].arrival_s))
      throw std::invalid_argument("em: request arrival times must be finite and non-decreasing");
    if (!std::isfinite(r.footprint_particles) || r.footprint_particles < 0.)
      throw std::invalid_argument("em: request footprint");
    if (r.particle_count) {
      if (!r.particles) throw std::invalid_argument("em: request " + std::to_string(k) + " has a null batch");
      for (std::uint64_t i = 0; i < r.particle_count; ++i)
        if (!can_transport(r.particles[i]))
          throw std::invalid_argument("em: request " + std::to_string(k) + " has a particle the device cannot transport");
    } else if (!can_transport(r.primary)) {
      throw std::invalid_argument("em: request " + std::to_string(k) + " has a primary the device cannot transport");
    }
    if (r.axis.grammage_g_per_cm2) {
      validate_axis(r.axis);
      axis_words += r.axis.support_count;
    }
    if (r.radio_observer_count && !slot_radio_)
      throw std::invalid_argument("em: request " + std::to_string(k) + " has radio observers, per-shower radio is off");
    std::memset(&results[k], 0, sizeof results[k]);
  }
  for (std::uint32_t slot = 0; slot < slots_; ++slot) slot_host_[slot].active = false;
  update_active_range();
  // every request's projection table, uploaded once
  C8MI_HIP(hipStreamSynchronize(stream_));
  DeviceBuffer<double> axes(axis_words ? axis_words : 1);
  std::vector<double const*> axis_of(count, nullptr);
  {
    std::vector<double> host(axis_words);
    std::uint64_t at = 0;
    for (std::uint64_t k = 0; k < count; ++k) {
      c8mi_em_shower_axis const& a = requests[k].axis;
      if (!a.grammage_g_per_cm2) continue;
      std::memcpy(host.data() + at, a.grammage_g_per_cm2, a.support_count * sizeof(double));
      axis_of[k] = axes.data() + at;
      at += a.support_count;
    }
    if (axis_words)
      C8MI_HIP(hipMemcpy(axes.data(), host.data(), axis_words * sizeof(double), hipMemcpyHostToDevice));
  }

  using clock = std::chrono::steady_clock;
  auto const t0 = clock::now();
  auto elapsed = [&] { return std::chrono::duration<double>(clock::now() - t0).count(); };
  std::vector<std::uint32_t> free_slots;
  for (std::uint32_t s = slots_; s-- > 0;) free_slots.push_back(s);
  std::vector<std::int64_t> owner(slots_, -1);
  std::vector<SlotInit> inits;
  std::vector<std::uint32_t> finished;
  std::uint64_t next = 0, active = 0;
  double footprint = 0.;
  while (next < count || active) {
    double const now = elapsed();
    inits.clear();
    while (next < count && !free_slots.empty() && (!options.realtime || requests[next].arrival_s <= now) &&
           (active == 0 || footprint + requests[next].footprint_particles <= options.footprint_budget_particles)) {
      c8mi_em_request const& r = requests[next];
      std::uint32_t const slot = free_slots.back();
      free_slots.pop_back();
      SlotInit init{};
      ShowerSlot& sh = init.shower;
      c8mi_em_shower_axis const* a = r.axis.grammage_g_per_cm2 ? &r.axis : nullptr;
      for (int d = 0; d < 3; ++d) {
        sh.axis_start[d] = a ? a->start_m[d] : config_.axis_start_m[d];
        sh.axis_direction[d] = a ? a->direction[d] : config_.axis_direction[d];
      }
      sh.axis_step = a ? a->step_m : config_.axis_step_m;
      sh.axis_grammage = a ? axis_of[next] : axis_.data();
      sh.axis_support_count = a ? a->support_count : config_.axis_support_count;
      sh.key = r.shower_id;
      init.first_history = r.first_history_id;
      init.slot = slot;
      inits.push_back(init);
      activate(slot, r.shower_id, r.first_history_id);
      if (slot_radio_) slot_radio_->set_slot(slot, r.radio_observers, r.radio_observer_count);
      if (r.particle_count) slot_enqueue(slot, r.particles, r.particle_count);
      else slot_enqueue(slot, &r.primary, 1);
      owner[slot] = static_cast<std::int64_t>(next);
      results[next].admitted_s = now;
      results[next].slot = slot;
      footprint += r.footprint_particles;
      ++active;
      ++next;
    }
    if (!inits.empty()) {
      start_slots(inits.data(), static_cast<std::uint32_t>(inits.size()));
      update_active_range();
    }
    if (active == 0) {  // idle until the next arrival
      double const wait = requests[next].arrival_s - elapsed();
      if (wait > 0.) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      continue;
    }
    c8mi_em_run_result round{};
    run_rounds_loop(1, round);
    totals.photon_fronts += round.photon_fronts;
    totals.lepton_fronts += round.lepton_fronts;
    totals.photon_steps += round.photon_steps;
    totals.lepton_steps += round.lepton_steps;
    totals.photon_interactions += round.photon_interactions;
    totals.lepton_interactions += round.lepton_interactions;
    totals.secondaries += round.secondaries;
    totals.max_front = std::max(totals.max_front, round.max_front);
    totals.photon_calls += round.photon_calls;
    totals.lepton_calls += round.lepton_calls;
    totals.bucketed_fronts += round.bucketed_fronts;
    totals.bucketed_particles += round.bucketed_particles;
    totals.below_minimum_batch_checkpoints += round.below_minimum_batch_checkpoints;
    totals.history_range_checkpoints += round.history_range_checkpoints;
    totals.reserved_history_ids += round.reserved_history_ids;
    finished.clear();
    for (std::uint32_t slot = 0; slot < slots_; ++slot) {
      if (owner[slot] < 0) continue;
      SlotHost const& h = slot_host_[slot];
      if (!(h.pending[0] || h.pending[1] || h.staged[0] || h.staged[1])) finished.push_back(slot);
    }
    if (finished.empty()) continue;
    // one gather and one read-back for every shower that ended this round
    std::uint32_t const n = static_cast<std::uint32_t>(finished.size());
    std::memcpy(gather_slots_host_.data(), finished.data(), n * sizeof(std::uint32_t));
    C8MI_HIP(hipMemcpyAsync(gather_slots_dev_.data(), gather_slots_host_.data(), n * sizeof(std::uint32_t),
                            hipMemcpyHostToDevice, stream_));
    launch_slot_gather(pool_arrays(), gather_slots_dev_.data(), n, gather_dev_.data(), stream_);
    C8MI_HIP(hipMemcpyAsync(gather_host_.data(), gather_dev_.data(),
                            static_cast<std::size_t>(n) * kGatherWords * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, stream_));
    C8MI_HIP(hipStreamSynchronize(stream_));
    double const done = elapsed();
    for (std::uint32_t k = 0; k < n; ++k) {
      std::uint32_t const slot = finished[k];
      unsigned long long const* g = gather_host_.data() + static_cast<std::size_t>(k) * kGatherWords;
      c8mi_em_request_result& out = results[owner[slot]];
      out.finished_s = done;
      out.photon_steps = g[kGatherStats + 0];
      out.lepton_steps = g[kGatherStats + 1];
      out.photon_interactions = g[kGatherStats + 2];
      out.lepton_interactions = g[kGatherStats + 3];
      out.secondaries = g[kGatherStats + 4];
      out.next_history_id = slot_host_[slot].call.counter;
      out.observations = g[kGatherRecords];
      out.fallbacks = g[kGatherRecords + 1];
      long long c[kCounters];
      for (int j = 0; j < kCounters; ++j) c[j] = static_cast<long long>(g[kGatherCounters + j]);
      decode_ledger(c, slot, out.ledger);
      c8mi_em_request const& req = requests[owner[slot]];
      if (slot_radio_ && req.radio_observer_count)
        slot_radio_->download(slot, req.radio_coreas, req.radio_zhs, &out.radio_counters);
      totals.observations += out.observations;
      totals.fallbacks += out.fallbacks;
      slot_host_[slot].active = false;
      footprint -= requests[owner[slot]].footprint_particles;
      owner[slot] = -1;
      free_slots.push_back(slot);
      --active;
    }
    update_active_range();
  }
  C8MI_HIP(hipStreamSynchronize(stream_));  // nothing may read `axes` after it is freed
}

void EmBackend::enable_timing(bool enable) {
  if (slot_radio_) slot_radio_->enable_kernel_timing(enable);
  timing_ = enable;
  ms_.fill(0.);
  launches_.fill(0);
}

void EmBackend::kernel_time(double ms[kFrontKernels], std::uint64_t launches[kFrontKernels]) const {
  for (int k = 0; k < kFrontKernels; ++k) {
    ms[k] = ms_[k];
    launches[k] = launches_[k];
  }
}

}  // namespace c8mi::em
