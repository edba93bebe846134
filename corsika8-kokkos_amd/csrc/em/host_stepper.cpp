#include "em/host_stepper.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "em/em_backend.hpp"
#include "em/steps.hpp"

namespace c8mi::em {

namespace {

template <class T>
std::vector<T> copy_of(T const* host, std::size_t count) {
  return count ? std::vector<T>(host, host + count) : std::vector<T>();
}

}  // namespace

HostStepper::HostStepper(c8mi_em_config const& config, c8mi_native_table_set const& tables,
                         c8mi_environment const& environment)
    : HostStepper(config, std::make_shared<HostTables const>(tables), environment) {}

HostStepper::HostStepper(c8mi_em_config const& config, std::shared_ptr<HostTables const> tables,
                         c8mi_environment const& environment)
    : config_(config), tables_(std::move(tables)) {
  if (!tables_) throw std::invalid_argument("em: null host tables");
  validate_config(config);
  validate_environment(environment);
  nodes_ = copy_of(environment.nodes, environment.node_count);
  children_ = copy_of(environment.children, environment.child_index_count);
  models_ = copy_of(environment.models, environment.model_count);
  axis_ = copy_of(config.axis_grammage_g_per_cm2, config.axis_support_count);
  std::uint64_t const bins = config.profile_bins ? config.profile_bins : 1u;
  counters_.assign(kCounters, 0);
  profile_.assign(4u * bins, 0);
  // the record arrays grow with use up to the configured capacities (a host
  // prefix writes a handful; the device-sized capacity would be mostly idle)
  observations_.resize(std::min<std::uint64_t>(config.observation_capacity, kInitialRecords));
  fallbacks_.resize(std::min<std::uint64_t>(config.fallback_capacity, kInitialRecords));

  EmContext& h = context_;
  fill_context(config, environment, h);
  h.tables = tables_->view();
  h.nodes = nodes_.data();
  h.children = children_.data();
  h.models = models_.data();
  // every Molière sample iterates to convergence here (the device's split
  // into a fast part and a tail computes the same iterates)
  h.moliere_fast_iterations = 100;
  if (config.moliere_enabled && config.electron_moliere_interpolation) {
    c8mi_moliere_interpolation const& t = *config.electron_moliere_interpolation;
    polynomials_ = copy_of(t.polynomials, sizeof(t.polynomials) / sizeof(t.polynomials[0]));
    guess_ = copy_of(t.initial_guess_delta, sizeof(t.initial_guess_delta) / sizeof(double));
    h.moliere_polynomials = polynomials_.data();
    h.moliere_guess = guess_.data();
  }
  for (int a = 0; a < 3; ++a) {
    shower_.axis_start[a] = config.axis_start_m[a];
    shower_.axis_direction[a] = config.axis_direction[a];
  }
  shower_.axis_step = config.axis_step_m;
  shower_.axis_grammage = axis_.data();
  shower_.axis_support_count = config.axis_support_count;
  h.showers = &shower_;
  h.shower_slots = 1;
  h.observations = observations_.data();
  h.fallbacks = fallbacks_.data();
  h.observation_capacity = observations_.size();
  h.fallback_capacity = fallbacks_.size();
  h.record_counts = record_counts_;
  h.first_interactions = &first_interaction_;
  h.first_candidates = &first_candidates_;
  h.counters = counters_.data();
  h.profile = profile_.data();
}

void HostStepper::begin_shower(std::uint64_t shower_id) {
  shower_.key = shower_id;
  std::fill(counters_.begin(), counters_.end(), 0);
  std::fill(profile_.begin(), profile_.end(), 0);
  record_counts_[0] = record_counts_[1] = 0;
  first_candidates_ = 0;
}

bool HostStepper::can_route(c8mi_particle_state const& p) const {
  if (!(p.pid == 22 || p.pid == 11 || p.pid == -11)) return false;
  if (!std::isfinite(p.energy_gev) || !(p.energy_gev > 0.) || !std::isfinite(p.weight) || p.weight < 0.)
    return false;
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(p.position_m[a]) || !std::isfinite(p.direction[a])) return false;
  return env::query(context_, p.position_m, p.direction, p.environment_node).status == env::kSuccess;
}

void HostStepper::book_injected(double energy_weight) {
  phys::Accumulator const acc{profile_.data(), counters_.data(), &context_, &shower_, 0};
  phys::ledger(acc, kInjected, energy_weight);
}

void HostStepper::reserve_records() {
  // one step records at most one observation or one fallback: room for two
  // more keeps every record up to the configured capacity
  auto grow = [](auto& records, std::uint64_t count, std::uint64_t limit) {
    if (count + 2 <= records.size() || records.size() >= limit) return false;
    records.resize(std::min<std::uint64_t>(limit, std::max<std::uint64_t>(2 * records.size(), count + 2)));
    return true;
  };
  if (grow(observations_, record_counts_[0], config_.observation_capacity)) {
    context_.observations = observations_.data();
    context_.observation_capacity = observations_.size();
  }
  if (grow(fallbacks_, record_counts_[1], config_.fallback_capacity)) {
    context_.fallbacks = fallbacks_.data();
    context_.fallback_capacity = fallbacks_.size();
  }
}

HostStep HostStepper::step(c8mi_particle_state const& particle, std::uint64_t first_history_id, bool moliere) {
  reserve_records();
  HostStep r{};
  phys::Accumulator const acc{profile_.data(), counters_.data(), &context_, &shower_, 0};
  phys::Particle p;
  static_cast<c8mi_particle_state&>(p) = particle;
  p.shower_key = shower_.key;
  c8mi_particle_state staging[2];
  phys::Outputs o;
  o.staging = staging;
  o.track = nullptr;
  o.track_stride = 1;
  o.kind[0] = o.kind[1] = kSlotNone;
  o.next = o.other = o.children = 0;
  o.step = o.interaction = o.tracked = false;
  EmContext const& c = context_;
  if (p.pid == 22) {
    phys::photon_step(c, acc, p, o);
  } else {
    o.track = r.track;
    phys::LeptonMove mv;
    if (phys::lepton_move(c, acc, p, mv)) {
      bool failed = false;
      if (c.moliere_enabled && moliere) {
        phys::MoliereOpen open;
        int const outcome = c.moliere_polynomials ? phys::moliere_scatter<true>(c, mv, open)
                                                  : phys::moliere_scatter<false>(c, mv, open);
        if (outcome == phys::kMoliereOpen) throw std::logic_error("em: host Moliere sample left open");
        failed = outcome == phys::kMoliereFailed;
      }
      phys::LeptonVertex vx;
      if (phys::lepton_finish(c, acc, [&] { return p; }, mv, failed, o, vx)) phys::lepton_vertex(c, acc, vx, o);
    }
    r.has_track = o.tracked;
  }
  r.step = o.step;
  r.interaction = o.interaction;
  r.children = o.children;
  for (int k = 0; k < 2; ++k) {
    if (o.kind[k] == kSlotNone) continue;
    c8mi_particle_state q = staging[k];
    if (o.kind[k] != kSlotNextContinue) q.history_id = first_history_id + q.history_id;
    q.pad_ = 0;
    q.pad2_ = 0;
    r.out[r.count++] = q;
    phys::patch_first_interaction(first_interaction_, first_candidates_, q);
  }
  return r;
}

}  // namespace c8mi::em
