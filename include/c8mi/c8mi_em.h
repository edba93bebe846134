/*
 * c8mi EM C ABI — MI355X replacement for the device side of CORSIKA8-kokkos's
 * electromagnetic backend (corsika/kokkos/em). Every entry point returns a
 * c8mi_status (0 = success, codes as in c8mi_radio.h); on failure
 * c8mi_last_error() returns the message for the calling thread.
 *
 * Device pointers are plain hipMalloc'd addresses (c8mi_device_malloc). All
 * launches are asynchronous on the given HIP stream (NULL = the null stream).
 */
#ifndef C8MI_EM_H
#define C8MI_EM_H

#include "c8mi_tables.h"
#include "c8mi_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* RandomNumberKey (corsika/kokkos/em/random/Philox.hpp:27-35), 48 bytes. */
typedef struct c8mi_random_key {
  uint64_t seed;
  uint64_t shower_id;
  uint64_t history_id;
  uint64_t step_id;
  uint32_t process_id;
  uint32_t pad_;
  uint64_t draw_id;
} c8mi_random_key;

/* Device copy of one format-v7 native table set: the storage the reference
 * allocates in BackendInstance::initialize from a ProposalNativeTableSet
 * (corsika/kokkos/em/tables/ProposalNativeTable.hpp:21-56,
 * BackendInstanceImplementation.inl). The host set is deep-copied; bounds of
 * every spline descriptor and index are checked first (invalid_argument). */
typedef struct c8mi_em_tables c8mi_em_tables;

int c8mi_em_tables_create(const c8mi_native_table_set* host_tables, c8mi_em_tables** out);
int c8mi_em_tables_destroy(c8mi_em_tables* tables);
/* The device view: same struct, device pointers (ProposalNativeDeviceView). */
int c8mi_em_tables_view(const c8mi_em_tables* tables, c8mi_native_table_set* device_view);
/* Device bytes held by the tables (proposal_native_table_bytes,
 * ProposalNativeTable.hpp:66-67). */
int c8mi_em_tables_bytes(const c8mi_em_tables* tables, uint64_t* bytes);

/* Batched PROPOSAL-native queries: one device thread per query, results as the
 * reference's query_proposal_native_* functions
 * (corsika/detail/kokkos/em/tables/ProposalNativeQueries.inl:744-1428) return
 * them (status, Newton/bisection iteration counts, value). Both arrays are
 * device memory of `count` entries. */
int c8mi_em_native_queries(const c8mi_em_tables* tables, const c8mi_native_query* device_queries,
                           uint64_t count, c8mi_native_result* device_results, void* stream);

/* Batched Philox4x32-10 draws (random_words / uniform_open01,
 * corsika/detail/kokkos/em/random/Philox.inl:23-57) for `count` device keys:
 * four words and the four open-interval uniforms of each key. Either output
 * may be NULL. */
int c8mi_em_random(const c8mi_random_key* device_keys, uint64_t count, uint32_t* device_words,
                   double* device_uniforms, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* C8MI_EM_H */
