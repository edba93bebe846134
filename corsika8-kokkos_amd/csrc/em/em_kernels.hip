// Batched table-query and Philox kernels of the EM path. The query kernel is
// the device form of the reference's execute_table_query over a batch
// (corsika/detail/kokkos/em/tables/TableQuery.inl:12-43, native kinds only);
// the random kernel evaluates random_words / uniform_open01
// (corsika/detail/kokkos/em/random/Philox.inl:23-57) for a list of keys.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "c8mi/c8mi_em.h"
#include "em/em_launch.hpp"
#include "em/native_tables.hpp"
#include "em/philox.hpp"
#include "runtime.hpp"

namespace c8mi::em {

namespace {

constexpr int kQueryBlock = 128;
constexpr int kRandomBlock = 256;

__global__ __launch_bounds__(kQueryBlock) void native_query_kernel(
    c8mi_native_table_set tables, c8mi_native_query const* __restrict__ queries,
    std::uint64_t count, c8mi_native_result* __restrict__ results) {
  std::uint64_t const i = static_cast<std::uint64_t>(blockIdx.x) * kQueryBlock + threadIdx.x;
  if (i >= count) return;
  c8mi_native_query const q = queries[i];
  QueryResult const r = run_query(tables, q);
  c8mi_native_result out;
  out.status = r.status;
  out.newton_iterations = r.newton;
  out.bisection_iterations = r.bisection;
  out.pad_ = 0;
  out.value = r.value;
  results[i] = out;
}

__global__ __launch_bounds__(kRandomBlock) void random_kernel(
    c8mi_random_key const* __restrict__ keys, std::uint64_t count,
    std::uint32_t* __restrict__ words, double* __restrict__ uniforms) {
  std::uint64_t const i = static_cast<std::uint64_t>(blockIdx.x) * kRandomBlock + threadIdx.x;
  if (i >= count) return;
  c8mi_random_key const k = keys[i];
  RandomKey const key{k.seed, k.shower_id, k.history_id, k.step_id, k.process_id, k.draw_id};
  std::uint32_t w[4];
  random_words(key, w);
  if (words) {
    uint4 packed;
    packed.x = w[0];
    packed.y = w[1];
    packed.z = w[2];
    packed.w = w[3];
    reinterpret_cast<uint4*>(words)[i] = packed;
  }
  if (uniforms) {
#pragma unroll
    for (int lane = 0; lane < 4; ++lane)
      uniforms[4 * i + lane] = (static_cast<double>(w[lane]) + 0.5) * (1. / 4294967296.);
  }
}

unsigned blocks_for(std::uint64_t count, int block) {
  std::uint64_t const blocks = (count + block - 1) / block;
  if (blocks > 0x7fffffffULL) throw std::length_error("batch exceeds the launch grid");
  return static_cast<unsigned>(blocks);
}

}  // namespace

void launch_native_queries(c8mi_native_table_set const& tables, c8mi_native_query const* queries,
                           std::uint64_t count, c8mi_native_result* results, hipStream_t stream) {
  if (count == 0) return;
  native_query_kernel<<<blocks_for(count, kQueryBlock), kQueryBlock, 0, stream>>>(tables, queries,
                                                                                 count, results);
  C8MI_HIP(hipGetLastError());
}

void launch_random(c8mi_random_key const* keys, std::uint64_t count, std::uint32_t* words,
                   double* uniforms, hipStream_t stream) {
  if (count == 0) return;
  random_kernel<<<blocks_for(count, kRandomBlock), kRandomBlock, 0, stream>>>(keys, count, words,
                                                                             uniforms);
  C8MI_HIP(hipGetLastError());
}

}  // namespace c8mi::em
