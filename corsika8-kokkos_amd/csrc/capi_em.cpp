// extern "C" EM entry points (include/c8mi/c8mi_em.h): device table storage,
// batched native-table queries and batched Philox draws.
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "c8mi/c8mi_em.h"
#include "capi_guard.hpp"
#include "em/device_tables.hpp"
#include "em/em_launch.hpp"
#include "runtime.hpp"

struct c8mi_em_tables {
  c8mi::em::DeviceTables tables;
  explicit c8mi_em_tables(c8mi_native_table_set const& host) : tables(host) {}
};

using c8mi::guarded;

extern "C" {

int c8mi_em_tables_create(c8mi_native_table_set const* host_tables, c8mi_em_tables** out) {
  return guarded([&] {
    if (host_tables == nullptr || out == nullptr) throw std::invalid_argument("null argument");
    *out = nullptr;
    *out = new c8mi_em_tables(*host_tables);
  });
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

}  // extern "C"
