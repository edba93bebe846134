// extern "C" EM entry points (include/c8mi/c8mi_em.h): device table storage,
// batched native-table queries and batched Philox draws.
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "c8mi/c8mi_em.h"
#include "capi_guard.hpp"
#include "capi_handles.hpp"
#include "em/device_tables.hpp"
#include "em/em_backend.hpp"
#include "em/em_launch.hpp"
#include "em/moliere_table.hpp"
#include "em/native_table_file.hpp"
#include "runtime.hpp"

struct c8mi_em_tables {
  c8mi::em::DeviceTables tables;
  explicit c8mi_em_tables(c8mi_native_table_set const& host) : tables(host) {}
};

struct c8mi_native_table_file {
  std::unique_ptr<c8mi::em::TableFile> file;
};

struct c8mi_em {
  c8mi::em::EmBackend backend;
  c8mi_em(c8mi_em_config const& config, c8mi_native_table_set const& tables,
          c8mi_environment const& environment, hipStream_t stream)
      : backend(config, tables, environment, stream) {}
};

using c8mi::guarded;

namespace {
void require_handle(void const* p) {
  if (p == nullptr) throw std::invalid_argument("null handle");
}
}  // namespace

extern "C" {

int c8mi_em_tables_create(c8mi_native_table_set const* host_tables, c8mi_em_tables** out) {
  return guarded([&] {
    if (host_tables == nullptr || out == nullptr) throw std::invalid_argument("null argument");
    *out = nullptr;
    *out = new c8mi_em_tables(*host_tables);
  });
}

int c8mi_sha256(uint8_t const* data, uint64_t size, uint8_t digest[32]) {
  return guarded([&] {
    if ((size && !data) || !digest) throw std::invalid_argument("null argument");
    c8mi::em::Digest const d = c8mi::em::sha256(data, size);
    std::memcpy(digest, d.data(), d.size());
  });
}

int c8mi_native_table_hash(c8mi_native_table_set const* set, c8mi_native_table_meta const* meta,
                           uint8_t digest[32]) {
  return guarded([&] {
    if (!set || !meta || !digest) throw std::invalid_argument("null argument");
    c8mi::em::Digest const d = c8mi::em::native_table_hash(*set, *meta);
    std::memcpy(digest, d.data(), d.size());
  });
}

int c8mi_native_table_write(char const* path, c8mi_native_table_set const* set, c8mi_native_table_meta const* meta) {
  return guarded([&] {
    if (!path || !set || !meta) throw std::invalid_argument("null argument");
    c8mi::em::write_native_table(path, *set, *meta);
  });
}

int c8mi_native_table_read(char const* path, c8mi_native_table_file** out) {
  return guarded([&] {
    if (!path || !out) throw std::invalid_argument("null argument");
    *out = nullptr;
    auto handle = std::make_unique<c8mi_native_table_file>();
    handle->file = c8mi::em::read_native_table(path);
    *out = handle.release();
  });
}

int c8mi_native_table_file_view(c8mi_native_table_file const* file, c8mi_native_table_set* set,
                                c8mi_native_table_meta* meta, uint8_t digest[32]) {
  return guarded([&] {
    require_handle(file);
    if (set) *set = file->file->set();
    if (meta) *meta = file->file->meta();
    if (digest) std::memcpy(digest, file->file->digest().data(), 32);
  });
}

int c8mi_native_table_file_destroy(c8mi_native_table_file* file) {
  return guarded([&] { delete file; });
}

int c8mi_em_tables_destroy(c8mi_em_tables* tables) {
  return guarded([&] { delete tables; });
}

int c8mi_em_tables_view(c8mi_em_tables const* tables, c8mi_native_table_set* device_view) {
  return guarded([&] {
    if (tables == nullptr || device_view == nullptr) throw std::invalid_argument("null argument");
    *device_view = tables->tables.view();
  });
}

int c8mi_em_tables_bytes(c8mi_em_tables const* tables, uint64_t* bytes) {
  return guarded([&] {
    if (tables == nullptr || bytes == nullptr) throw std::invalid_argument("null argument");
    *bytes = tables->tables.bytes();
  });
}

int c8mi_em_native_queries(c8mi_em_tables const* tables, c8mi_native_query const* device_queries,
                           uint64_t count, c8mi_native_result* device_results, void* stream) {
  return guarded([&] {
    if (tables == nullptr) throw std::invalid_argument("null tables");
    if (count > 0 && (device_queries == nullptr || device_results == nullptr))
      throw std::invalid_argument("null query or result array");
    c8mi::em::launch_native_queries(tables->tables.view(), device_queries, count, device_results,
                                    static_cast<hipStream_t>(stream));
  });
}

int c8mi_em_random(c8mi_random_key const* device_keys, uint64_t count, uint32_t* device_words,
                   double* device_uniforms, void* stream) {
  return guarded([&] {
    if (count > 0 && device_keys == nullptr) throw std::invalid_argument("null key array");
    c8mi::em::launch_random(device_keys, count, device_words, device_uniforms,
                            static_cast<hipStream_t>(stream));
  });
}

int c8mi_em_create(c8mi_em_config const* config, c8mi_native_table_set const* tables,
                   c8mi_environment const* environment, void* stream, c8mi_em** out) {
  return guarded([&] {
    if (!config || !tables || !environment || !out) throw std::invalid_argument("null argument");
    *out = nullptr;
    *out = new c8mi_em(*config, *tables, *environment, static_cast<hipStream_t>(stream));
  });
}

int c8mi_em_destroy(c8mi_em* em) {
  return guarded([&] { delete em; });
}

int c8mi_em_begin_shower(c8mi_em* em, uint64_t shower_id, uint64_t first_history_id) {
  return guarded([&] {
    require_handle(em);
    em->backend.begin_shower(shower_id, first_history_id);
  });
}

int c8mi_em_can_transport(c8mi_em const* em, c8mi_particle_state const* particle, int* out) {
  return guarded([&] {
    require_handle(em);
    if (!particle || !out) throw std::invalid_argument("null argument");
    *out = em->backend.can_transport(*particle) ? 1 : 0;
  });
}

int c8mi_em_enqueue(c8mi_em* em, c8mi_particle_state const* particles, uint64_t count) {
  return guarded([&] {
    require_handle(em);
    em->backend.enqueue(particles, count);
  });
}

int c8mi_em_run_photon_wavefront(c8mi_em* em, c8mi_particle_state const* particles, uint64_t count,
                                 uint64_t first_secondary_history_id, uint64_t maximum_wavefronts,
                                 uint64_t minimum_resident_batch, c8mi_em_wavefront_result* out) {
  return guarded([&] {
    require_handle(em);
    if (!out) throw std::invalid_argument("null result");
    if (count && !particles) throw std::invalid_argument("null particle array");
    for (uint64_t i = 0; i < count; ++i)
      if (particles[i].pid != 22) throw std::invalid_argument("runPhotonWavefront takes photons only");
    em->backend.run_call(true, particles, count, first_secondary_history_id, maximum_wavefronts, 0,
                         minimum_resident_batch, *out);
  });
}

int c8mi_em_run_lepton_wavefront(c8mi_em* em, c8mi_particle_state const* particles, uint64_t count,
                                 uint64_t first_secondary_history_id, uint64_t maximum_wavefronts,
                                 uint64_t secondary_history_id_limit_exclusive,
                                 uint64_t minimum_resident_batch, c8mi_em_wavefront_result* out) {
  return guarded([&] {
    require_handle(em);
    if (!out) throw std::invalid_argument("null result");
    if (count && !particles) throw std::invalid_argument("null particle array");
    for (uint64_t i = 0; i < count; ++i)
      if (particles[i].pid != 11 && particles[i].pid != -11)
        throw std::invalid_argument("runLeptonWavefront takes e+- only");
    em->backend.run_call(false, particles, count, first_secondary_history_id, maximum_wavefronts,
                         secondary_history_id_limit_exclusive, minimum_resident_batch, *out);
  });
}

int c8mi_em_download_remaining(c8mi_em* em, c8mi_particle_state* out, uint64_t capacity, uint64_t* count) {
  return guarded([&] {
    require_handle(em);
    auto const& r = em->backend.remaining();
    uint64_t const n = r.size() < capacity ? r.size() : capacity;
    if (n && !out) throw std::invalid_argument("null output array");
    for (uint64_t i = 0; i < n; ++i) out[i] = r[i];
    if (count) *count = r.size();
  });
}

int c8mi_em_get_limits(c8mi_em const* em, c8mi_em_limits* out) {
  return guarded([&] {
    require_handle(em);
    if (!out) throw std::invalid_argument("null output");
    em->backend.limits(*out);
  });
}

int c8mi_em_get_statistics(c8mi_em const* em, c8mi_em_statistics* out) {
  return guarded([&] {
    require_handle(em);
    if (!out) throw std::invalid_argument("null output");
    em->backend.statistics(*out);
  });
}

int c8mi_em_run(c8mi_em* em, uint64_t maximum_rounds, c8mi_em_run_result* out) {
  return guarded([&] {
    require_handle(em);
    if (!out) throw std::invalid_argument("null result");
    em->backend.run(maximum_rounds, *out);
  });
}

int c8mi_em_pending_counts(c8mi_em const* em, uint64_t* photons, uint64_t* leptons) {
  return guarded([&] {
    require_handle(em);
    if (!photons || !leptons) throw std::invalid_argument("null argument");
    uint64_t p = 0, l = 0;
    em->backend.pending(p, l);
    *photons = p;
    *leptons = l;
  });
}

int c8mi_em_download_profile(c8mi_em* em, double* photons, double* electrons, double* positrons,
                             double* energy_loss, c8mi_em_ledger* ledger) {
  return guarded([&] {
    require_handle(em);
    em->backend.download_profile(photons, electrons, positrons, energy_loss, ledger);
  });
}

int c8mi_em_download_observations(c8mi_em* em, c8mi_particle_state* out, uint64_t capacity,
                                  uint64_t* count) {
  return guarded([&] {
    require_handle(em);
    uint64_t const n = em->backend.download_observations(out, capacity);
    if (count) *count = n;
  });
}

int c8mi_em_download_fallbacks(c8mi_em* em, c8mi_em_fallback* out, uint64_t capacity, uint64_t* count) {
  return guarded([&] {
    require_handle(em);
    uint64_t const n = em->backend.download_fallbacks(out, capacity);
    if (count) *count = n;
  });
}

int c8mi_em_download_first_interaction(c8mi_em* em, c8mi_em_first_interaction* out, int* captured) {
  return c8mi_em_slot_download_first_interaction(em, 0, out, captured);
}

int c8mi_em_slot_download_first_interaction(c8mi_em* em, uint32_t slot, c8mi_em_first_interaction* out,
                                            int* captured) {
  return guarded([&] {
    require_handle(em);
    if (!out || !captured) throw std::invalid_argument("null argument");
    *captured = em->backend.slot_download_first_interaction(slot, *out) ? 1 : 0;
  });
}

int c8mi_em_slot_begin(c8mi_em* em, uint32_t slot, uint64_t shower_id, uint64_t first_history_id,
                       c8mi_em_shower_axis const* axis) {
  return guarded([&] {
    require_handle(em);
    em->backend.slot_begin(slot, shower_id, first_history_id, axis);
  });
}

int c8mi_em_slot_enqueue(c8mi_em* em, uint32_t slot, c8mi_particle_state const* particles, uint64_t count) {
  return guarded([&] {
    require_handle(em);
    em->backend.slot_enqueue(slot, particles, count);
  });
}

int c8mi_em_run_rounds(c8mi_em* em, uint64_t maximum_rounds, c8mi_em_run_result* out) {
  return guarded([&] {
    require_handle(em);
    if (!out) throw std::invalid_argument("null result");
    em->backend.run(maximum_rounds, *out);
  });
}

int c8mi_em_slot_state_get(c8mi_em const* em, uint32_t slot, c8mi_em_slot_state* out) {
  return guarded([&] {
    require_handle(em);
    if (!out) throw std::invalid_argument("null result");
    em->backend.slot_state(slot, *out);
  });
}

int c8mi_em_slot_download_profile(c8mi_em* em, uint32_t slot, double* photons, double* electrons,
                                  double* positrons, double* energy_loss, c8mi_em_ledger* ledger) {
  return guarded([&] {
    require_handle(em);
    em->backend.slot_download_profile(slot, photons, electrons, positrons, energy_loss, ledger);
  });
}

int c8mi_em_slot_download_observations(c8mi_em* em, uint32_t slot, c8mi_particle_state* out,
                                       uint64_t capacity, uint64_t* count) {
  return guarded([&] {
    require_handle(em);
    uint64_t const n = em->backend.slot_download_records(slot, 0, out, sizeof(c8mi_particle_state), capacity);
    if (count) *count = n;
  });
}

int c8mi_em_slot_download_fallbacks(c8mi_em* em, uint32_t slot, c8mi_em_fallback* out, uint64_t capacity,
                                    uint64_t* count) {
  return guarded([&] {
    require_handle(em);
    uint64_t const n = em->backend.slot_download_records(slot, 1, out, sizeof(c8mi_em_fallback), capacity);
    if (count) *count = n;
  });
}

int c8mi_em_slot_end(c8mi_em* em, uint32_t slot) {
  return guarded([&] {
    require_handle(em);
    em->backend.slot_end(slot);
  });
}

int c8mi_em_slot_pending(c8mi_em const* em, uint32_t slot, uint64_t* photons, uint64_t* leptons) {
  return guarded([&] {
    require_handle(em);
    if (!photons || !leptons) throw std::invalid_argument("null argument");
    uint64_t p = 0, l = 0;
    em->backend.slot_pending(slot, p, l);
    *photons = p;
    *leptons = l;
  });
}

int c8mi_em_serve(c8mi_em* em, c8mi_em_request const* requests, uint64_t count,
                  c8mi_em_serve_options const* options, c8mi_em_request_result* results,
                  c8mi_em_run_result* totals) {
  return guarded([&] {
    require_handle(em);
    if (!options || !totals) throw std::invalid_argument("null argument");
    em->backend.serve(requests, count, *options, results, *totals);
  });
}

int c8mi_em_moliere_interpolation_build(c8mi_moliere_snapshot const* snapshot,
                                        c8mi_moliere_interpolation* out) {
  return guarded([&] {
    if (!snapshot || !out) throw std::invalid_argument("null argument");
    c8mi::em::build_moliere_table(*snapshot, *out);
  });
}

int c8mi_em_moliere_interpolation_read(char const* path, c8mi_moliere_snapshot const* snapshot,
                                       c8mi_moliere_interpolation* out) {
  return guarded([&] {
    if (!path || !snapshot || !out) throw std::invalid_argument("null argument");
    if (!c8mi::em::read_moliere_table(path, *snapshot, *out))
      throw std::invalid_argument(std::string("no valid Moliere cache for this snapshot in ") + path);
  });
}

int c8mi_em_moliere_interpolation_write(char const* path, c8mi_moliere_snapshot const* snapshot,
                                        c8mi_moliere_interpolation const* table) {
  return guarded([&] {
    if (!path || !snapshot || !table) throw std::invalid_argument("null argument");
    if (!c8mi::em::valid_moliere_table(*table)) throw std::invalid_argument("Moliere table with non-finite entries");
    c8mi::em::write_moliere_table(path, *snapshot, *table);
  });
}

int c8mi_em_attach_radio(c8mi_em* em, c8mi_radio* radio) {
  return guarded([&] {
    require_handle(em);
    em->backend.attach_radio(radio ? &radio->accumulator : nullptr);
  });
}

int c8mi_em_enable_slot_radio(c8mi_em* em, c8mi_radio_config const* config, uint32_t observers_per_slot,
                              uint32_t bins_per_observer) {
  return guarded([&] {
    require_handle(em);
    if (!config) throw std::invalid_argument("null radio configuration");
    em->backend.enable_slot_radio(*config, observers_per_slot, bins_per_observer);
  });
}

int c8mi_em_slot_set_radio(c8mi_em* em, uint32_t slot, c8mi_radio_observer const* observers, uint64_t count) {
  return guarded([&] {
    require_handle(em);
    em->backend.slot_set_radio(slot, observers, count);
  });
}

int c8mi_em_slot_download_radio(c8mi_em* em, uint32_t slot, int64_t* coreas, int64_t* zhs,
                                c8mi_radio_counters* counters) {
  return guarded([&] {
    require_handle(em);
    em->backend.slot_download_radio(slot, coreas, zhs, counters);
  });
}

int c8mi_em_slot_radio_kernel_time(c8mi_em* em, double* ms, uint64_t* launches) {
  return guarded([&] {
    require_handle(em);
    if (!ms || !launches) throw std::invalid_argument("null argument");
    em->backend.slot_radio_time(*ms, *launches);
  });
}

int c8mi_em_enable_kernel_timing(c8mi_em* em, int enable) {
  return guarded([&] {
    require_handle(em);
    em->backend.enable_timing(enable != 0);
  });
}

int c8mi_em_kernel_time(c8mi_em const* em, double ms[6], uint64_t launches[6]) {
  return guarded([&] {
    require_handle(em);
    if (!ms || !launches) throw std::invalid_argument("null argument");
    em->backend.kernel_time(ms, launches);
  });
}

}  // extern "C"
