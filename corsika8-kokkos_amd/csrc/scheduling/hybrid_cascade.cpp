#include "scheduling/hybrid_cascade.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>

#include "runtime.hpp"

namespace c8mi::scheduling {

namespace {

// FallbackReason::MoliereSamplingFailed (FallbackRequest.hpp:22-65).
constexpr std::int32_t kMoliereSamplingFailed = 21;

// A status of the EM / radio C ABI back into the exception it stands for.
void check(int status) {
  if (status == kOk) return;
  std::string const message = last_error();
  switch (status) {
    case kInvalidArgument: throw std::invalid_argument(message);
    case kLogicError: throw std::logic_error(message);
    case kLengthError: throw std::length_error(message);
    case kDeviceError: throw DeviceError(message);
    case kOverflowError: throw OverflowError(message);
    default: throw std::runtime_error(message);
  }
}

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point start) {
  return std::chrono::duration<double>(Clock::now() - start).count();
}

bool same_transport_state(c8mi_particle_state const& a, c8mi_particle_state const& b) {
  if (a.pid != b.pid || a.environment_node != b.environment_node || a.medium_id != b.medium_id ||
      a.energy_gev != b.energy_gev || a.time_s != b.time_s || a.weight != b.weight)
    return false;
  for (int k = 0; k < 3; ++k)
    if (a.position_m[k] != b.position_m[k] || a.direction[k] != b.direction[k]) return false;
  return true;
}

}  // namespace

c8mi_em_hybrid_config resolve_config(c8mi_em_hybrid_config const* config) {
  c8mi_em_hybrid_config c{};
  if (config) c = *config;
  if (!c.minimum_batch_size) c.minimum_batch_size = 16384;
  if (!c.maximum_expansion_rounds) c.maximum_expansion_rounds = 8;
  if (!c.photon_wavefronts) c.photon_wavefronts = 16;
  if (!c.lepton_wavefronts) c.lepton_wavefronts = 1024;
  if (!c.lepton_reservation_multiplier) c.lepton_reservation_multiplier = 64;
  if (!c.lepton_reservation_chunk) c.lepton_reservation_chunk = std::uint64_t{1} << 20;
  return c;
}

HybridCascade::HybridCascade(c8mi_em* em, em::HostStepper& stepper, c8mi_em_hybrid_config const* config)
    : em_(em), stepper_(stepper), config_(resolve_config(config)) {
  if (em_) {
    check(c8mi_em_get_limits(em_, &limits_));
  } else {
    limits_.maximum_resident_photon_batch = limits_.maximum_resident_lepton_batch =
        limits_.maximum_resident_input_batch = std::numeric_limits<std::uint64_t>::max();
  }
}

void HybridCascade::start(c8mi_particle_state const* primaries, std::uint64_t count, std::uint64_t shower_id,
                          std::uint64_t first_history_id) {
  if (count && !primaries) throw std::invalid_argument("hybrid: null primary array");
  if (first_history_id == 0) throw std::invalid_argument("hybrid: history IDs start at 1");
  stack_.clear();
  staged_.clear();
  expansion_steps_.clear();
  fallback_steps_.clear();
  fallback_cursor_ = 0;
  expanded_.clear();
  tracks_.clear();
  force_small_ = false;
  consecutive_ = 0;
  counter_ = first_history_id;
  stop_ = false;
  host_seconds_ = device_seconds_ = 0.;
  r_ = c8mi_em_hybrid_result{};
  stepper_.begin_shower(shower_id);
  if (em_) check(c8mi_em_begin_shower(em_, shower_id, first_history_id));
  for (std::uint64_t i = 0; i < count; ++i) {
    stepper_.book_injected(primaries[i].energy_gev * primaries[i].weight);
    stack_.push_back(primaries[i]);
  }
}

HybridCascade::Route HybridCascade::route(c8mi_particle_state const& p) {
  // EmRouter::canRoute (EmRouter.inl:64-105): a step returned for a host
  // expansion or a fallback is declined once
  auto const expansion = expansion_steps_.find({p.history_id, p.step_id});
  if (expansion != expansion_steps_.end()) {
    expansion_steps_.erase(expansion);
    return Route::kHostStep;
  }
  auto const fallback = fallback_steps_.find({p.history_id, p.step_id});
  if (fallback != fallback_steps_.end()) {
    fallback_steps_.erase(fallback);
    return Route::kFallbackStep;
  }
  return stepper_.can_route(p) ? Route::kDevice : Route::kHostStep;
}

void HybridCascade::return_fallbacks() {
  // EmRouter::returnFallbacks (EmRouter.inl:709-800): the shower's new device
  // fallback records; a Molière sampling failure goes back to the host stack
  // for a host step without Molière scattering (the reference hands it to
  // C8's scalar transport), in (history, step) order so that the host's
  // history IDs do not depend on the order the device appended the records
  std::uint64_t total = 0;
  check(c8mi_em_download_fallbacks(em_, nullptr, 0, &total));
  if (total <= fallback_cursor_) return;
  std::uint64_t const stored = std::min<std::uint64_t>(total, stepper_.config().fallback_capacity);
  fallbacks_.resize(stored);
  std::uint64_t n = 0;
  if (stored) check(c8mi_em_download_fallbacks(em_, fallbacks_.data(), fallbacks_.size(), &n));
  std::vector<c8mi_particle_state> returned;
  for (std::uint64_t i = fallback_cursor_; i < stored; ++i) {
    c8mi_em_fallback const& f = fallbacks_[i];
    if (f.reason == kMoliereSamplingFailed) returned.push_back(f.particle);
    else ++r_.fallbacks_kept;
  }
  if (total > std::max(stored, fallback_cursor_)) r_.fallbacks_kept += total - std::max(stored, fallback_cursor_);
  std::sort(returned.begin(), returned.end(), [](c8mi_particle_state const& a, c8mi_particle_state const& b) {
    return a.history_id != b.history_id ? a.history_id < b.history_id : a.step_id < b.step_id;
  });
  for (auto const& p : returned) {
    if (!fallback_steps_.insert({p.history_id, p.step_id}).second)
      throw std::runtime_error("hybrid: duplicate host fallback for one transport step");
    stack_.push_back(p);
    stepper_.book_injected(p.energy_gev * p.weight);
    ++r_.fallbacks_returned;
  }
  fallback_cursor_ = total;
}

void HybridCascade::stage(c8mi_particle_state const& p) {
  auto const previous = expanded_.find(p.history_id);
  if (previous != expanded_.end() && same_transport_state(previous->second, p)) force_small_ = true;
  staged_.push_back(p);
  ++r_.particles_staged;
}

void HybridCascade::host_step(c8mi_particle_state const& p, bool moliere) {
  auto const t0 = Clock::now();
  em::HostStep const s = stepper_.step(p, counter_, moliere);
  counter_ += s.children;
  ++r_.cpu_particles;
  if (s.step) {
    if (p.pid == 22) ++r_.cpu_photon_steps;
    else ++r_.cpu_lepton_steps;
  }
  if (s.interaction) ++r_.cpu_interactions;
  r_.cpu_secondaries += s.children;
  if (s.has_track) {
    tracks_.insert(tracks_.end(), s.track, s.track + 9);
    ++r_.cpu_tracks;
  }
  for (std::uint32_t k = 0; k < s.count; ++k) stack_.push_back(s.out[k]);
  host_seconds_ += seconds_since(t0);
}

std::uint64_t HybridCascade::device_pending() const {
  if (!em_) return 0;
  std::uint64_t photons = 0, leptons = 0;
  check(c8mi_em_pending_counts(em_, &photons, &leptons));
  return photons + leptons;
}

bool HybridCascade::ready() const {
  return force_small_ || device_pending() != 0 || staged_.size() >= config_.minimum_batch_size;
}

bool HybridCascade::pending() const { return !staged_.empty() || device_pending() != 0; }

void HybridCascade::flush_tracks() {
  if (!radio_ || tracks_.empty()) return;
  std::size_t const n = tracks_.size() / 9;
  records_.assign(n, c8mi_lepton_transport_record{});
  for (std::size_t i = 0; i < n; ++i) {
    double const* t = tracks_.data() + 9 * i;
    c8mi_lepton_transport_record& r = records_[i];
    for (int a = 0; a < 3; ++a) {
      r.start.position_m[a] = t[a];
      r.end.position_m[a] = t[3 + a];
    }
    r.start.time_s = t[6];
    r.end.time_s = t[7];
    r.start.pid = std::signbit(t[8]) ? 11 : -11;  // the weight is signed by charge, negative for e-
    r.start.weight = std::fabs(t[8]);
    r.end.pid = r.start.pid;
    r.end.weight = r.start.weight;
  }
  check(c8mi_radio_enqueue_host_tracks(radio_, records_.data(), n));
  tracks_.clear();
}

void HybridCascade::device_call(bool lepton, std::vector<c8mi_particle_state> const& host) {
  auto const t0 = Clock::now();
  std::uint64_t pending_photons = 0, pending_leptons = 0;
  check(c8mi_em_pending_counts(em_, &pending_photons, &pending_leptons));
  std::uint64_t const n = host.size() + (lepton ? pending_leptons : pending_photons);
  r_.maximum_input_batch = std::max(r_.maximum_input_batch, n);
  c8mi_em_wavefront_result w{};
  std::uint64_t const first = counter_;
  std::uint64_t count;
  if (!lepton) {
    // reservePhotonSecondaryRange (EmRouter.inl:640-656)
    if (n > std::numeric_limits<std::uint64_t>::max() / 2 / config_.photon_wavefronts)
      throw OverflowError("hybrid: photon history reservation overflow");
    count = n * 2 * config_.photon_wavefronts;
    counter_ += count;
    check(c8mi_em_run_photon_wavefront(em_, host.data(), host.size(), first, config_.photon_wavefronts,
                                       config_.minimum_batch_size, &w));
  } else {
    // reserveResidentSecondaryRange (EmRouter.inl:658-679)
    if (n > std::numeric_limits<std::uint64_t>::max() / 3) throw OverflowError("hybrid: lepton reservation overflow");
    std::uint64_t const minimum = 3 * n;
    count = minimum;
    if (minimum <= std::numeric_limits<std::uint64_t>::max() / config_.lepton_reservation_multiplier)
      count = std::max(std::min(config_.lepton_reservation_chunk, minimum * config_.lepton_reservation_multiplier),
                       minimum);
    counter_ += count;
    check(c8mi_em_run_lepton_wavefront(em_, host.data(), host.size(), first, config_.lepton_wavefronts,
                                       first + count, config_.minimum_batch_size, &w));
  }
  r_.reserved_history_ids += count;
  for (auto const& p : host) stepper_.book_injected(-(p.energy_gev * p.weight));
  if (lepton) {
    ++r_.lepton_calls;
    r_.device_lepton_fronts += w.wavefronts;
    r_.device_lepton_steps += w.transport_records;
  } else {
    ++r_.photon_calls;
    r_.device_photon_fronts += w.wavefronts;
    r_.device_photon_steps += w.transport_records;
  }
  r_.device_interactions += w.interactions;
  r_.device_secondaries += w.secondaries;
  r_.bucketed_fronts += w.bucketed_fronts;
  r_.below_minimum_batch_checkpoints += w.below_minimum_batch_checkpoint;
  r_.history_range_checkpoints += w.history_range_exhausted;
  // remainingPhotons_ / remainingLeptons_ join the staged batch
  remaining_.resize(w.remaining);
  std::uint64_t copied = 0;
  if (w.remaining) check(c8mi_em_download_remaining(em_, remaining_.data(), remaining_.size(), &copied));
  for (auto const& p : remaining_) {
    staged_.push_back(p);
    stepper_.book_injected(p.energy_gev * p.weight);
  }
  return_fallbacks();
  device_seconds_ += seconds_since(t0);
}

void HybridCascade::advance() {
  flush_tracks();
  std::uint64_t const pending = device_pending();
  if (staged_.empty() && pending == 0) return;
  std::uint64_t const minimum = config_.minimum_batch_size;
  if (staged_.size() < minimum && !force_small_ && pending == 0 && consecutive_ < config_.maximum_expansion_rounds) {
    // scalar expansion (EmRouter.inl:260-293): one host step for each
    expanded_.clear();
    for (auto const& p : staged_) {
      if (!expansion_steps_.insert({p.history_id, p.step_id}).second)
        throw std::runtime_error("hybrid: duplicate host wavefront-expansion step");
      stack_.push_back(p);
      expanded_.emplace(p.history_id, p);
    }
    r_.particles_returned_for_expansion += staged_.size();
    staged_.clear();
    force_small_ = false;
    ++consecutive_;
    ++r_.small_batch_expansions;
    return;
  }
  if (staged_.size() < minimum) {
    if (force_small_) ++r_.stalled_boundary_flushes;
    else if (pending == 0 && consecutive_ >= config_.maximum_expansion_rounds) ++r_.scalar_expansion_budget_flushes;
  }
  force_small_ = false;
  consecutive_ = 0;
  expanded_.clear();
  if (prefix_only_) {
    stop_ = true;
    return;
  }
  std::vector<c8mi_particle_state> current;
  current.swap(staged_);
  if (current.size() > limits_.maximum_resident_input_batch) {
    staged_.assign(current.begin() + static_cast<std::ptrdiff_t>(limits_.maximum_resident_input_batch), current.end());
    current.resize(limits_.maximum_resident_input_batch);
    ++r_.input_batch_splits;
  }
  ++r_.wavefronts;
  std::vector<c8mi_particle_state> photons, leptons;
  for (auto const& p : current) (p.pid == 22 ? photons : leptons).push_back(p);
  std::uint64_t pending_photons = 0, pending_leptons = 0;
  check(c8mi_em_pending_counts(em_, &pending_photons, &pending_leptons));
  std::uint64_t const photon_chunk = std::min(pending_photons, limits_.maximum_resident_photon_batch);
  std::uint64_t const host_photons = limits_.maximum_resident_photon_batch - photon_chunk;
  if (photons.size() > host_photons) {
    staged_.insert(staged_.end(), photons.begin() + static_cast<std::ptrdiff_t>(host_photons), photons.end());
    photons.resize(host_photons);
    ++r_.input_batch_splits;
  }
  if (photons.size() + photon_chunk) device_call(false, photons);
  check(c8mi_em_pending_counts(em_, &pending_photons, &pending_leptons));
  if (leptons.size() + pending_leptons) {
    std::uint64_t const lepton_chunk = std::min(pending_leptons, limits_.maximum_resident_lepton_batch);
    std::uint64_t const host_leptons = limits_.maximum_resident_lepton_batch - lepton_chunk;
    if (leptons.size() > host_leptons) {
      staged_.insert(staged_.end(), leptons.begin() + static_cast<std::ptrdiff_t>(host_leptons), leptons.end());
      leptons.resize(host_leptons);
      ++r_.input_batch_splits;
    }
    device_call(true, leptons);
  }
}

void HybridCascade::cascade() {
  while (!stop_ && (!stack_.empty() || pending())) {
    while (!stop_ && !stack_.empty()) {
      c8mi_particle_state const p = stack_.back();
      stack_.pop_back();
      Route const r = route(p);
      if (r == Route::kDevice) {
        stage(p);
        while (!stop_ && ready()) advance();
        continue;
      }
      host_step(p, r != Route::kFallbackStep);
    }
    if (!stop_ && pending()) advance();
  }
}

c8mi_em_hybrid_result HybridCascade::run(c8mi_particle_state const* primaries, std::uint64_t count,
                                         std::uint64_t shower_id, std::uint64_t first_history_id) {
  if (!em_) throw std::logic_error("hybrid: no EM handle");
  prefix_only_ = false;
  start(primaries, count, shower_id, first_history_id);
  cascade();
  flush_tracks();
  if (radio_) check(c8mi_radio_flush_host_tracks(radio_));
  r_.next_history_id = counter_;
  r_.host_seconds = host_seconds_;
  r_.device_seconds = device_seconds_;
  return r_;
}

c8mi_em_hybrid_result HybridCascade::prefix(c8mi_particle_state const* primaries, std::uint64_t count,
                                            std::uint64_t shower_id, std::uint64_t first_history_id,
                                            std::vector<c8mi_particle_state>& batch, std::vector<double>& tracks,
                                            std::uint64_t& next_history_id) {
  c8mi_em* const em = em_;
  em_ = nullptr;  // the prefix needs no device
  prefix_only_ = true;
  try {
    start(primaries, count, shower_id, first_history_id);
    cascade();
  } catch (...) {
    em_ = em;
    prefix_only_ = false;
    throw;
  }
  em_ = em;
  prefix_only_ = false;
  if (!stack_.empty()) throw std::logic_error("hybrid: particles left on the host stack at the handoff");
  for (auto const& p : staged_) stepper_.book_injected(-(p.energy_gev * p.weight));
  batch = staged_;
  tracks = tracks_;
  next_history_id = counter_;
  r_.next_history_id = counter_;
  r_.host_seconds = host_seconds_;
  return r_;
}

}  // namespace c8mi::scheduling
