// Wavefront kernels of the EM backend. One front of one species is three
// launches on the backend's stream:
//   1. step     one thread per queued particle: the whole step (selection,
//               transport, vertex, final state) of PhotonFront.inl /
//               LeptonFront.inl fused into one kernel; up to two output
//               particles go to a per-source staging pair; profile and ledger
//               contributions accumulate in block-private LDS copies that are
//               added to the device totals once per block; per-block output
//               counts are reduced for the scan.
//   2. scan     one workgroup: exclusive scan of the per-block counts (next,
//               other species, children) and the front totals.
//   3. scatter  stable compaction: block-local prefix of the per-source counts
//               plus the block offset gives every output its queue slot, so the
//               order (and the history IDs first + child offset + k) follows
//               the source order exactly, as the reference's stable scans do.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "em/em_device.hpp"
#include "em/em_front.hpp"
#include "em/steps.hpp"
#include "runtime.hpp"

namespace c8mi::em {

namespace {

constexpr int kStepBlock = 256;
constexpr int kScanBlock = 1024;

struct BlockCounts {
  std::uint32_t next, other, children, steps, interactions, pad[3];
};

__device__ __forceinline__ std::uint32_t wave_sum(std::uint32_t v) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, 64);
  return v;
}

template <bool Photon>
__global__ __launch_bounds__(kStepBlock) void step_kernel(EmContext const* __restrict__ context,
                                                          ParticleQueue in, std::uint64_t count,
                                                          c8mi_particle_state* __restrict__ staging,
                                                          std::uint32_t* __restrict__ packed,
                                                          BlockCounts* __restrict__ block_counts) {
  extern __shared__ long long lds[];
  EmContext const& c = *context;
  std::uint32_t const bins = c.profile_bins;
  long long* hist = lds;
  long long* counters = lds + 4u * bins;
  for (std::uint32_t k = threadIdx.x; k < 4u * bins + kCounters; k += kStepBlock) lds[k] = 0;
  __syncthreads();
  phys::Accumulator const acc{hist, counters, context};

  std::uint64_t const i = static_cast<std::uint64_t>(blockIdx.x) * kStepBlock + threadIdx.x;
  std::uint32_t word = 0;
  std::uint32_t next = 0, other = 0, children = 0, steps = 0, interactions = 0;
  if (i < count) {
    phys::Particle const p = load_particle(in, i);
    phys::Outputs o;
    o.kind[0] = o.kind[1] = kSlotNone;
    o.next = o.other = o.children = 0;
    o.step = o.interaction = false;
    if constexpr (Photon) phys::photon_step(c, acc, p, o);
    else phys::lepton_step(c, acc, p, o);
    if (o.kind[0] != kSlotNone) staging[2 * i] = o.slot[0];
    if (o.kind[1] != kSlotNone) staging[2 * i + 1] = o.slot[1];
    word = pack_counts(o.next, o.other, o.children, o.kind[0], o.kind[1]);
    packed[i] = word;
    next = o.next;
    other = o.other;
    children = o.children;
    steps = o.step ? 1u : 0u;
    interactions = o.interaction ? 1u : 0u;
  }
  // per-block totals
  __shared__ std::uint32_t partial[kStepBlock / 64][5];
  std::uint32_t const wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  std::uint32_t const s0 = wave_sum(next), s1 = wave_sum(other), s2 = wave_sum(children);
  std::uint32_t const s3 = wave_sum(steps), s4 = wave_sum(interactions);
  if (lane == 0) {
    partial[wave][0] = s0;
    partial[wave][1] = s1;
    partial[wave][2] = s2;
    partial[wave][3] = s3;
    partial[wave][4] = s4;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    BlockCounts b{0, 0, 0, 0, 0, {0, 0, 0}};
    for (int w = 0; w < kStepBlock / 64; ++w) {
      b.next += partial[w][0];
      b.other += partial[w][1];
      b.children += partial[w][2];
      b.steps += partial[w][3];
      b.interactions += partial[w][4];
    }
    block_counts[blockIdx.x] = b;
  }
  // flush the block's profile and ledger contributions
  for (std::uint32_t k = threadIdx.x; k < 4u * bins; k += kStepBlock) {
    long long const v = hist[k];
    if (v != 0) atomicAdd(reinterpret_cast<unsigned long long*>(&c.profile[k]), static_cast<unsigned long long>(v));
  }
  for (std::uint32_t k = threadIdx.x; k < kCounters; k += kStepBlock) {
    long long const v = counters[k];
    if (v != 0) atomicAdd(reinterpret_cast<unsigned long long*>(&c.counters[k]), static_cast<unsigned long long>(v));
  }
}

__global__ __launch_bounds__(kScanBlock) void scan_kernel(BlockCounts const* __restrict__ counts,
                                                          std::uint32_t blocks,
                                                          unsigned long long* __restrict__ offsets,
                                                          FrontTotals* __restrict__ totals) {
  __shared__ unsigned long long carry[5];
  __shared__ unsigned long long wave_sums[kScanBlock / 64][5];
  if (threadIdx.x < 5) carry[threadIdx.x] = 0;
  __syncthreads();
  std::uint32_t const wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  for (std::uint32_t base = 0; base < blocks; base += kScanBlock) {
    std::uint32_t const b = base + threadIdx.x;
    unsigned long long v[5] = {0, 0, 0, 0, 0};
    if (b < blocks) {
      BlockCounts const bc = counts[b];
      v[0] = bc.next;
      v[1] = bc.other;
      v[2] = bc.children;
      v[3] = bc.steps;
      v[4] = bc.interactions;
    }
    // inclusive wave scan
    unsigned long long inc[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      unsigned long long x = v[f];
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        unsigned long long const y = __shfl_up(x, d, 64);
        if (static_cast<int>(lane) >= d) x += y;
      }
      inc[f] = x;
    }
    if (lane == 63)
      for (int f = 0; f < 5; ++f) wave_sums[wave][f] = inc[f];
    __syncthreads();
    unsigned long long before[5] = {0, 0, 0, 0, 0};
    for (std::uint32_t w = 0; w < wave; ++w)
      for (int f = 0; f < 5; ++f) before[f] += wave_sums[w][f];
    if (b < blocks)
      for (int f = 0; f < 3; ++f) offsets[3 * static_cast<std::uint64_t>(b) + f] = carry[f] + before[f] + inc[f] - v[f];
    __syncthreads();
    if (threadIdx.x == kScanBlock - 1)
      for (int f = 0; f < 5; ++f) carry[f] += before[f] + inc[f];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    FrontTotals t{};
    t.next = carry[0];
    t.other = carry[1];
    t.children = carry[2];
    t.steps = carry[3];
    t.interactions = carry[4];
    *totals = t;
  }
}

__global__ __launch_bounds__(kStepBlock) void scatter_kernel(
    std::uint32_t const* __restrict__ packed, std::uint64_t count,
    c8mi_particle_state const* __restrict__ staging, unsigned long long const* __restrict__ offsets,
    ParticleQueue next_queue, ParticleQueue other_queue, std::uint64_t other_base,
    std::uint64_t first_history) {
  __shared__ std::uint32_t wave_sums[kStepBlock / 64][3];
  std::uint64_t const i = static_cast<std::uint64_t>(blockIdx.x) * kStepBlock + threadIdx.x;
  std::uint32_t const word = i < count ? packed[i] : 0u;
  std::uint32_t const v[3] = {word & 3u, (word >> 2) & 3u, (word >> 4) & 3u};
  std::uint32_t const wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  std::uint32_t inc[3];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    std::uint32_t x = v[f];
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      std::uint32_t const y = __shfl_up(x, d, 64);
      if (static_cast<int>(lane) >= d) x += y;
    }
    inc[f] = x;
  }
  if (lane == 63)
    for (int f = 0; f < 3; ++f) wave_sums[wave][f] = inc[f];
  __syncthreads();
  if (i >= count) return;
  std::uint64_t pos[3];
  for (int f = 0; f < 3; ++f) {
    std::uint32_t before = 0;
    for (std::uint32_t w = 0; w < wave; ++w) before += wave_sums[w][f];
    pos[f] = offsets[3 * static_cast<std::uint64_t>(blockIdx.x) + f] + before + inc[f] - v[f];
  }
  std::uint64_t next_pos = pos[0], other_pos = other_base + pos[1];
  std::uint64_t const child_base = first_history + pos[2];
  for (int s = 0; s < 2; ++s) {
    std::uint32_t const kind = (word >> (6 + 2 * s)) & 3u;
    if (kind == kSlotNone) continue;
    c8mi_particle_state p = staging[2 * i + s];
    if (kind != kSlotNextContinue) p.history_id = child_base + p.history_id;
    if (kind == kSlotOtherChild) store_particle(other_queue, other_pos++, p);
    else store_particle(next_queue, next_pos++, p);
  }
}

unsigned grid_for(std::uint64_t n, int block) {
  std::uint64_t const g = (n + block - 1) / block;
  if (g > 0x7fffffffULL) throw std::length_error("front exceeds the launch grid");
  return static_cast<unsigned>(g);
}

}  // namespace

std::size_t step_lds_bytes(std::uint32_t bins) {
  return (4u * static_cast<std::size_t>(bins) + kCounters) * sizeof(long long);
}

std::size_t block_count_bytes(std::uint64_t n) { return grid_for(n, kStepBlock) * sizeof(BlockCounts); }

std::size_t offset_bytes(std::uint64_t n) { return grid_for(n, kStepBlock) * 3 * sizeof(unsigned long long); }

void launch_step(bool photon, EmContext const* context, std::uint32_t bins, ParticleQueue const& in,
                 std::uint64_t count, c8mi_particle_state* staging, std::uint32_t* packed,
                 void* block_counts, hipStream_t stream) {
  if (count == 0) return;
  unsigned const grid = grid_for(count, kStepBlock);
  std::size_t const lds = step_lds_bytes(bins);
  auto* bc = static_cast<BlockCounts*>(block_counts);
  if (photon)
    step_kernel<true><<<grid, kStepBlock, lds, stream>>>(context, in, count, staging, packed, bc);
  else
    step_kernel<false><<<grid, kStepBlock, lds, stream>>>(context, in, count, staging, packed, bc);
  C8MI_HIP(hipGetLastError());
}

void launch_scan(void const* block_counts, std::uint64_t count, unsigned long long* offsets,
                 FrontTotals* totals, hipStream_t stream) {
  std::uint32_t const blocks = count == 0 ? 0u : grid_for(count, kStepBlock);
  scan_kernel<<<1, kScanBlock, 0, stream>>>(static_cast<BlockCounts const*>(block_counts), blocks,
                                             offsets, totals);
  C8MI_HIP(hipGetLastError());
}

void launch_scatter(std::uint32_t const* packed, std::uint64_t count, c8mi_particle_state const* staging,
                    unsigned long long const* offsets, ParticleQueue const& next_queue,
                    ParticleQueue const& other_queue, std::uint64_t other_base,
                    std::uint64_t first_history, hipStream_t stream) {
  if (count == 0) return;
  scatter_kernel<<<grid_for(count, kStepBlock), kStepBlock, 0, stream>>>(
      packed, count, staging, offsets, next_queue, other_queue, other_base, first_history);
  C8MI_HIP(hipGetLastError());
}

}  // namespace c8mi::em
