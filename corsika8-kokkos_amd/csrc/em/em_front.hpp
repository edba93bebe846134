// Host launchers of the EM front kernels (em_front.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "em/em_device.hpp"

namespace c8mi::em {

std::size_t step_lds_bytes(std::uint32_t bins);
std::size_t block_count_bytes(std::uint64_t n);
std::size_t offset_bytes(std::uint64_t n);

void launch_step(bool photon, EmContext const* context, std::uint32_t bins, ParticleQueue const& in,
                 std::uint64_t count, c8mi_particle_state* staging, std::uint32_t* packed,
                 void* block_counts, hipStream_t stream);

void launch_scan(void const* block_counts, std::uint64_t count, unsigned long long* offsets,
                 FrontTotals* totals, hipStream_t stream);

void launch_scatter(std::uint32_t const* packed, std::uint64_t count, c8mi_particle_state const* staging,
                    unsigned long long const* offsets, ParticleQueue const& next_queue,
                    ParticleQueue const& other_queue, std::uint64_t other_base,
                    std::uint64_t first_history, hipStream_t stream);

}  // namespace c8mi::em
