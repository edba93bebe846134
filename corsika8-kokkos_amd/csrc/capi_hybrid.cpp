// extern "C" host handoff entry points (include/c8mi/c8mi_hybrid.h): the host
// scalar stepper and the cascade driver over the EM C ABI.
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <thread>
#include <vector>

#include "c8mi/c8mi_hybrid.h"
#include "capi_guard.hpp"
#include "em/em_backend.hpp"
#include "em/host_stepper.hpp"
#include "runtime.hpp"
#include "scheduling/hybrid_cascade.hpp"

struct c8mi_em_scalar {
  c8mi::em::HostStepper stepper;
  c8mi_em_scalar(c8mi_em_config const& config, c8mi_native_table_set const& tables,
                 c8mi_environment const& environment)
      : stepper(config, tables, environment) {}
};

struct c8mi_em_hybrid {
  c8mi::scheduling::HybridCascade cascade;
  c8mi_em_hybrid(c8mi_em* em, c8mi::em::HostStepper& stepper, c8mi_em_hybrid_config const* config)
      : cascade(em, stepper, config) {}
};

using c8mi::guarded;

namespace {
void require_handle(void const* p) {
  if (p == nullptr) throw std::invalid_argument("null handle");
}
}  // namespace

extern "C" {

int c8mi_em_scalar_create(c8mi_em_config const* config, c8mi_native_table_set const* tables,
                          c8mi_environment const* environment, c8mi_em_scalar** out) {
  return guarded([&] {
    if (!config || !tables || !environment || !out) throw std::invalid_argument("null argument");
    *out = nullptr;
    *out = new c8mi_em_scalar(*config, *tables, *environment);
  });
}

int c8mi_em_scalar_destroy(c8mi_em_scalar* scalar) {
  return guarded([&] { delete scalar; });
}

int c8mi_em_scalar_begin_shower(c8mi_em_scalar* scalar, uint64_t shower_id) {
  return guarded([&] {
    require_handle(scalar);
    scalar->stepper.begin_shower(shower_id);
  });
}

int c8mi_em_scalar_step(c8mi_em_scalar* scalar, c8mi_particle_state const* particle, uint64_t first_history_id,
                        c8mi_particle_state out[2], uint32_t* count, uint32_t* children, double track[9],
                        int* has_track) {
  return guarded([&] {
    require_handle(scalar);
    if (!particle || !out || !count) throw std::invalid_argument("null argument");
    c8mi::em::HostStep const s = scalar->stepper.step(*particle, first_history_id);
    for (std::uint32_t k = 0; k < s.count; ++k) out[k] = s.out[k];
    *count = s.count;
    if (children) *children = s.children;
    if (has_track) *has_track = s.has_track ? 1 : 0;
    if (track && s.has_track) std::memcpy(track, s.track, sizeof s.track);
  });
}

int c8mi_em_scalar_download_profile(c8mi_em_scalar const* scalar, double* photons, double* electrons,
                                    double* positrons, double* energy_loss, c8mi_em_ledger* ledger) {
  return guarded([&] {
    require_handle(scalar);
    auto const& s = scalar->stepper;
    std::uint32_t const bins = s.bins();
    double* outs[4] = {photons, electrons, positrons, energy_loss};
    for (int h = 0; h < 4; ++h) {
      if (!outs[h]) continue;
      double const scale = h == 3 ? s.config().profile_energy_scale : s.config().profile_weight_scale;
      for (std::uint32_t k = 0; k < bins; ++k) outs[h][k] = static_cast<double>(s.profile()[h * bins + k]) / scale;
    }
    if (ledger) c8mi::em::ledger_from_counters(s.counters(), s.config().profile_energy_scale, *ledger);
  });
}

int c8mi_em_scalar_download_observations(c8mi_em_scalar const* scalar, c8mi_particle_state* out, uint64_t capacity,
                                         uint64_t* count) {
  return guarded([&] {
    require_handle(scalar);
    auto const& s = scalar->stepper;
    std::uint64_t const stored = std::min<std::uint64_t>(s.observation_count(), s.config().observation_capacity);
    std::uint64_t const n = std::min(stored, capacity);
    if (n && !out) throw std::invalid_argument("null output array");
    for (std::uint64_t i = 0; i < n; ++i) out[i] = s.observations()[i];
    if (count) *count = s.observation_count();
  });
}

int c8mi_em_scalar_download_fallbacks(c8mi_em_scalar const* scalar, c8mi_em_fallback* out, uint64_t capacity,
                                      uint64_t* count) {
  return guarded([&] {
    require_handle(scalar);
    auto const& s = scalar->stepper;
    std::uint64_t const stored = std::min<std::uint64_t>(s.fallback_count(), s.config().fallback_capacity);
    std::uint64_t const n = std::min(stored, capacity);
    if (n && !out) throw std::invalid_argument("null output array");
    for (std::uint64_t i = 0; i < n; ++i) out[i] = s.fallbacks()[i];
    if (count) *count = s.fallback_count();
  });
}

int c8mi_em_scalar_download_first_interaction(c8mi_em_scalar const* scalar, c8mi_em_first_interaction* out,
                                              int* captured) {
  return guarded([&] {
    require_handle(scalar);
    if (!out || !captured) throw std::invalid_argument("null argument");
    *captured = scalar->stepper.first_interaction(*out) ? 1 : 0;
  });
}

int c8mi_em_hybrid_create(c8mi_em* em, c8mi_em_scalar* scalar, c8mi_em_hybrid_config const* config,
                          c8mi_em_hybrid** out) {
  return guarded([&] {
    if (!em || !scalar || !out) throw std::invalid_argument("null argument");
    *out = nullptr;
    *out = new c8mi_em_hybrid(em, scalar->stepper, config);
  });
}

int c8mi_em_hybrid_destroy(c8mi_em_hybrid* hybrid) {
  return guarded([&] { delete hybrid; });
}

int c8mi_em_hybrid_attach_radio(c8mi_em_hybrid* hybrid, c8mi_radio* radio) {
  return guarded([&] {
    require_handle(hybrid);
    hybrid->cascade.attach_radio(radio);
  });
}

int c8mi_em_hybrid_run(c8mi_em_hybrid* hybrid, c8mi_particle_state const* primaries, uint64_t count,
                       uint64_t shower_id, uint64_t first_history_id, c8mi_em_hybrid_result* out) {
  return guarded([&] {
    require_handle(hybrid);
    c8mi_em_hybrid_result const r = hybrid->cascade.run(primaries, count, shower_id, first_history_id);
    if (out) *out = r;
  });
}

int c8mi_em_hybrid_prefix(c8mi_em_scalar* scalar, c8mi_em_hybrid_config const* config,
                          c8mi_particle_state const* primaries, uint64_t count, uint64_t shower_id,
                          uint64_t first_history_id, c8mi_particle_state* staged, uint64_t capacity,
                          uint64_t* staged_count, uint64_t* next_history_id, double* tracks, uint64_t track_capacity,
                          uint64_t* track_count, c8mi_em_hybrid_result* out) {
  return guarded([&] {
    require_handle(scalar);
    c8mi::scheduling::HybridCascade cascade(nullptr, scalar->stepper, config);
    std::vector<c8mi_particle_state> batch;
    std::vector<double> host_tracks;
    std::uint64_t next = 0;
    c8mi_em_hybrid_result const r =
        cascade.prefix(primaries, count, shower_id, first_history_id, batch, host_tracks, next);
    std::uint64_t const n = std::min<std::uint64_t>(batch.size(), capacity);
    if (n && !staged) throw std::invalid_argument("null staged array");
    for (std::uint64_t i = 0; i < n; ++i) staged[i] = batch[i];
    if (staged_count) *staged_count = batch.size();
    if (next_history_id) *next_history_id = next;
    std::uint64_t const t = std::min<std::uint64_t>(host_tracks.size() / 9, track_capacity);
    if (t && !tracks) throw std::invalid_argument("null track array");
    if (t) std::memcpy(tracks, host_tracks.data(), t * 9 * sizeof(double));
    if (track_count) *track_count = host_tracks.size() / 9;
    if (out) *out = r;
  });
}

int c8mi_em_hybrid_prefixes(c8mi_em_config const* config, c8mi_native_table_set const* tables,
                            c8mi_environment const* environment, c8mi_em_hybrid_config const* hybrid,
                            c8mi_particle_state const* primaries, uint64_t count, uint64_t const* shower_ids,
                            uint64_t first_history_id, uint32_t threads, c8mi_particle_state* staged,
                            uint64_t capacity, uint64_t* offsets, uint64_t* next_history_ids,
                            c8mi_em_ledger* ledgers, double* histograms, c8mi_em_hybrid_result* totals) {
  return guarded([&] {
    if (!config || !tables || !environment || !offsets || !next_history_ids)
      throw std::invalid_argument("null argument");
    if (count && (!primaries || !shower_ids)) throw std::invalid_argument("null primary or shower ID array");
    std::uint32_t const workers = std::max<std::uint32_t>(1, std::min<std::uint64_t>(threads ? threads : 1, count));
    std::uint32_t const bins = config->profile_bins;
    std::vector<std::vector<c8mi_particle_state>> batches(count);
    std::vector<c8mi_em_hybrid_result> results(count);
    std::vector<std::exception_ptr> errors(workers);
    // one stepper per worker over one shared host copy of the tables; the
    // first is built here so that argument errors surface at once, the others
    // by their workers
    auto const host_tables = std::make_shared<c8mi::em::HostTables const>(*tables);
    std::vector<std::unique_ptr<c8mi::em::HostStepper>> steppers(workers);
    steppers[0] = std::make_unique<c8mi::em::HostStepper>(*config, host_tables, *environment);
    auto work = [&](std::uint32_t w) {
      try {
        if (!steppers[w]) steppers[w] = std::make_unique<c8mi::em::HostStepper>(*config, host_tables, *environment);
        c8mi::em::HostStepper& stepper = *steppers[w];
        c8mi::scheduling::HybridCascade cascade(nullptr, stepper, hybrid);
        std::vector<double> tracks;
        for (std::uint64_t k = w; k < count; k += workers) {
          results[k] = cascade.prefix(&primaries[k], 1, shower_ids[k], first_history_id, batches[k], tracks,
                                      next_history_ids[k]);
          if (ledgers) c8mi::em::ledger_from_counters(stepper.counters(), config->profile_energy_scale, ledgers[k]);
          if (histograms) {
            double* out = histograms + k * 4u * bins;
            for (int h = 0; h < 4; ++h) {
              double const scale = h == 3 ? config->profile_energy_scale : config->profile_weight_scale;
              for (std::uint32_t b = 0; b < bins; ++b)
                out[h * bins + b] = static_cast<double>(stepper.profile()[h * bins + b]) / scale;
            }
          }
        }
      } catch (...) {
        errors[w] = std::current_exception();
      }
    };
    std::vector<std::thread> pool;
    for (std::uint32_t w = 1; w < workers; ++w) pool.emplace_back(work, w);
    work(0);
    for (auto& t : pool) t.join();
    for (auto const& e : errors)
      if (e) std::rethrow_exception(e);
    std::uint64_t total = 0;
    for (std::uint64_t k = 0; k < count; ++k) {
      offsets[k] = total;
      total += batches[k].size();
    }
    offsets[count] = total;
    if (total > capacity) throw std::length_error("prefixes: the batches exceed the staged capacity");
    if (total && !staged) throw std::invalid_argument("null staged array");
    for (std::uint64_t k = 0; k < count; ++k)
      if (!batches[k].empty())
        std::memcpy(staged + offsets[k], batches[k].data(), batches[k].size() * sizeof(c8mi_particle_state));
    if (totals) {
      c8mi_em_hybrid_result t{};
      std::uint64_t* sum = &t.cpu_particles;
      constexpr std::size_t words = offsetof(c8mi_em_hybrid_result, next_history_id) / sizeof(std::uint64_t);
      for (auto const& r : results) {
        std::uint64_t const* add = &r.cpu_particles;
        for (std::size_t i = 0; i < words; ++i) sum[i] += add[i];
        t.maximum_input_batch = std::max(t.maximum_input_batch, r.maximum_input_batch);
        t.host_seconds += r.host_seconds;
      }
      *totals = t;
    }
  });
}

}  // extern "C"
