// Host-side launchers of the EM kernels (defined in the .hip files of this
// directory), callable from plain C++ translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "c8mi/c8mi_em.h"

namespace c8mi::em {

void launch_native_queries(c8mi_native_table_set const& tables, c8mi_native_query const* queries,
                           std::uint64_t count, c8mi_native_result* results, hipStream_t stream);

void launch_random(c8mi_random_key const* keys, std::uint64_t count, std::uint32_t* words,
                   double* uniforms, hipStream_t stream);

}  // namespace c8mi::em
