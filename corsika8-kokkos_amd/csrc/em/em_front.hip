// Wavefront kernels of the EM backend. One front of one species is three
// launches on the backend's stream:
//   1. step     one thread per queued particle: the whole step (selection,
//               transport, vertex, final state) of PhotonFront.inl /
//               LeptonFront.inl fused into one kernel; up to two output
//               particles go to a per-source staging pair; profile and ledger
//               contributions accumulate in block-private LDS copies (one for
//               the shower of the block's first particle, one for that of its
//               last) that are added to the shower's device totals once per
//               block; per-block output counts and per-(slot, run) output
//               totals are reduced for the scan.
//   2. scan     one workgroup: exclusive scan of the per-block counts, and per
//               slot the position and history offsets of every run's outputs.
//   3. scatter  stable compaction: the block prefix of the per-source counts
//               plus the block offset gives the source's global output rank;
//               the (slot, run) offset turns it into its rank inside its own
//               shower, so each shower's queue segment, order and history IDs
//               (first + child rank + k) are those of the shower run alone, as
//               the reference's stable scans give them.
// Queues are sorted by slot within each run (the previous front's output is
// one run, the other species' secondaries and host-staged particles each
// append one).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "em/em_device.hpp"
#include "em/em_front.hpp"
#include "em/steps.hpp"
#include "runtime.hpp"

namespace c8mi::em {

namespace {

constexpr int kStepBlock = 256;
constexpr int kScanBlock = 1024;
constexpr std::uint32_t kNoKey = 0xffffffffu;
constexpr int kRunShift = 20;

struct BlockCounts {
  std::uint32_t next, other, children, steps, interactions, pad[3];
};

__device__ __forceinline__ std::uint32_t run_of(FrontArgs const& f, std::uint64_t i) {
  std::uint32_t r = 0;
  while (r + 1 < f.runs && i >= f.run_end[r]) ++r;
  return r;
}

template <class T>
__device__ __forceinline__ T wave_inclusive(T x, std::uint32_t lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T const y = __shfl_up(x, d, 64);
    if (static_cast<int>(lane) >= d) x += y;
  }
  return x;
}

template <bool Photon>
__global__ __launch_bounds__(kStepBlock) void step_kernel(EmContext const* __restrict__ context, FrontArgs f) {
  extern __shared__ long long lds[];
  __shared__ std::uint32_t keys[kStepBlock];
  __shared__ std::uint32_t wave_sums[kStepBlock / 64][kShowerFields];
  EmContext const& c = *context;
  std::uint32_t const bins = c.profile_bins;
  std::uint32_t const stride = 4u * bins + kCounters;
  std::uint64_t const first = static_cast<std::uint64_t>(blockIdx.x) * kStepBlock;
  std::uint64_t const last = (first + kStepBlock < f.count ? first + kStepBlock : f.count) - 1;
  std::uint32_t const sh0 = f.in.shower[first], sh1 = f.in.shower[last];
  std::uint32_t const copies = sh0 == sh1 ? 1u : 2u;
  for (std::uint32_t k = threadIdx.x; k < copies * stride; k += kStepBlock) lds[k] = 0;
  __syncthreads();

  std::uint64_t const i = first + threadIdx.x;
  std::uint32_t v[kShowerFields] = {0, 0, 0, 0, 0};
  std::uint32_t key = kNoKey;
  if (i < f.count) {
    std::uint32_t const s = f.in.shower[i];
    long long* hist;
    if (s == sh0) hist = lds;
    else if (s == sh1) hist = lds + stride;
    else hist = nullptr;
    phys::Accumulator const acc{
        hist ? hist : c.profile + static_cast<std::uint64_t>(s) * 4u * bins,
        hist ? hist + 4u * bins : c.counters + static_cast<std::uint64_t>(s) * kCounters, context,
        &c.showers[s], s};
    phys::Particle p;
    static_cast<c8mi_particle_state&>(p) = load_particle(f.in, i);
    p.shower_key = c.showers[s].key;
    phys::Outputs o;
    o.kind[0] = o.kind[1] = kSlotNone;
    o.next = o.other = o.children = 0;
    o.step = o.interaction = false;
    if constexpr (Photon) phys::photon_step(c, acc, p, o);
    else phys::lepton_step(c, acc, p, o);
    if (o.kind[0] != kSlotNone) f.staging[2 * i] = o.slot[0];
    if (o.kind[1] != kSlotNone) f.staging[2 * i + 1] = o.slot[1];
    f.packed[i] = pack_counts(o.next, o.other, o.children, o.kind[0], o.kind[1], s);
    v[0] = o.next;
    v[1] = o.other;
    v[2] = o.children;
    v[3] = o.step ? 1u : 0u;
    v[4] = o.interaction ? 1u : 0u;
    key = (run_of(f, i) << kRunShift) | s;
  }
  keys[threadIdx.x] = key;
  std::uint32_t const wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  std::uint32_t inc[kShowerFields];
#pragma unroll
  for (int k = 0; k < kShowerFields; ++k) inc[k] = wave_inclusive(v[k], lane);
  if (lane == 63)
    for (int k = 0; k < kShowerFields; ++k) wave_sums[wave][k] = inc[k];
  __syncthreads();

  if (threadIdx.x == 0) {
    std::uint32_t t[kShowerFields] = {0, 0, 0, 0, 0};
    for (int w = 0; w < kStepBlock / 64; ++w)
      for (int k = 0; k < kShowerFields; ++k) t[k] += wave_sums[w][k];
    static_cast<BlockCounts*>(f.block_counts)[blockIdx.x] = BlockCounts{t[0], t[1], t[2], t[3], t[4], {0, 0, 0}};
  }
  // Per-(slot, run) totals: within the block the keys are non-decreasing
  // segments; the first source of a segment subtracts its exclusive block
  // prefix and the last adds its inclusive prefix (arithmetic modulo 2^64).
  if (key != kNoKey) {
    std::uint32_t const s = key & ((1u << kRunShift) - 1u), r = key >> kRunShift;
    unsigned long long* t =
        f.shower_totals + (static_cast<std::uint64_t>(s) * kMaxRuns + r) * kShowerFields;
    bool const start = threadIdx.x == 0 || keys[threadIdx.x - 1] != key;
    bool const end = threadIdx.x == kStepBlock - 1 || keys[threadIdx.x + 1] != key;
#pragma unroll
    for (int k = 0; k < kShowerFields; ++k) {
      std::uint32_t before = 0;
      for (std::uint32_t w = 0; w < wave; ++w) before += wave_sums[w][k];
      std::uint32_t const incl = before + inc[k], excl = incl - v[k];
      if (start && excl != 0) atomicAdd(&t[k], 0ULL - excl);
      if (end && incl != 0) atomicAdd(&t[k], static_cast<unsigned long long>(incl));
    }
  }
  // flush the block's profile and ledger contributions to the showers' totals
  for (std::uint32_t copy = 0; copy < copies; ++copy) {
    std::uint32_t const s = copy ? sh1 : sh0;
    long long const* src = lds + copy * stride;
    long long* hist = c.profile + static_cast<std::uint64_t>(s) * 4u * bins;
    long long* counters = c.counters + static_cast<std::uint64_t>(s) * kCounters;
    for (std::uint32_t k = threadIdx.x; k < stride; k += kStepBlock) {
      long long const value = src[k];
      if (value == 0) continue;
      long long* dst = k < 4u * bins ? hist + k : counters + (k - 4u * bins);
      atomicAdd(reinterpret_cast<unsigned long long*>(dst), static_cast<unsigned long long>(value));
    }
  }
}

__global__ __launch_bounds__(kScanBlock) void scan_kernel(FrontArgs f, std::uint32_t blocks, int photon) {
  // ---- per-block offsets of the three output streams and the front totals ----
  __shared__ unsigned long long carry[kShowerFields];
  __shared__ unsigned long long wave_sums[kScanBlock / 64][kShowerFields];
  if (threadIdx.x < kShowerFields) carry[threadIdx.x] = 0;
  __syncthreads();
  std::uint32_t const wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  auto const* counts = static_cast<BlockCounts const*>(f.block_counts);
  for (std::uint32_t base = 0; base < blocks; base += kScanBlock) {
    std::uint32_t const b = base + threadIdx.x;
    unsigned long long v[kShowerFields] = {0, 0, 0, 0, 0};
    if (b < blocks) {
      BlockCounts const bc = counts[b];
      v[0] = bc.next;
      v[1] = bc.other;
      v[2] = bc.children;
      v[3] = bc.steps;
      v[4] = bc.interactions;
    }
    unsigned long long inc[kShowerFields];
#pragma unroll
    for (int k = 0; k < kShowerFields; ++k) inc[k] = wave_inclusive(v[k], lane);
    if (lane == 63)
      for (int k = 0; k < kShowerFields; ++k) wave_sums[wave][k] = inc[k];
    __syncthreads();
    unsigned long long before[kShowerFields] = {0, 0, 0, 0, 0};
    for (std::uint32_t w = 0; w < wave; ++w)
      for (int k = 0; k < kShowerFields; ++k) before[k] += wave_sums[w][k];
    if (b < blocks)
      for (int k = 0; k < 3; ++k)
        f.offsets[3 * static_cast<std::uint64_t>(b) + k] = carry[k] + before[k] + inc[k] - v[k];
    __syncthreads();
    if (threadIdx.x == kScanBlock - 1)
      for (int k = 0; k < kShowerFields; ++k) carry[k] += before[k] + inc[k];
    __syncthreads();
  }

  // ---- per slot: offsets that turn a global output rank into the rank inside
  // the shower, new queue counts, history counters and statistics. One
  // exclusive scan over the slots per (run, stream). ----
  __shared__ unsigned long long slot_sums[2][kScanBlock / 64];
  std::uint32_t const s = threadIdx.x;
  bool const mine = s < f.slots;
  unsigned long long* row = f.shower_totals + static_cast<std::uint64_t>(s) * kMaxRuns * kShowerFields;
  // partial[r][k] = this slot's outputs in earlier runs - the rank offset of
  // the slot inside run r - the outputs of all slots in earlier runs
  unsigned long long partial[kMaxRuns][3];
  unsigned long long base[3] = {0, 0, 0}, within[3] = {0, 0, 0}, run_start[3] = {0, 0, 0};
  int parity = 0;
#pragma unroll
  for (int r = 0; r < kMaxRuns; ++r) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      partial[r][k] = 0;
      if (r >= static_cast<int>(f.runs)) continue;  // uniform
      unsigned long long const x = mine ? row[r * kShowerFields + k] : 0ULL;
      unsigned long long const inc = wave_inclusive(x, lane);
      if (lane == 63) slot_sums[parity][wave] = inc;
      __syncthreads();
      unsigned long long before = 0, all = 0;
      for (std::uint32_t w = 0; w < kScanBlock / 64; ++w) {
        unsigned long long const y = slot_sums[parity][w];
        before += w < wave ? y : 0ULL;
        all += y;
      }
      parity ^= 1;  // the next scan writes the other buffer; the one after is past a barrier
      unsigned long long const excl = before + inc - x;
      partial[r][k] = within[k] - excl - run_start[k];
      base[k] += excl;
      within[k] += x;
      run_start[k] += all;
    }
  }
  if (mine) {
    unsigned long long const history = f.history[s];
    base[2] = history;
#pragma unroll
    for (int r = 0; r < kMaxRuns; ++r)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (r < static_cast<int>(f.runs))
          f.adjust[(static_cast<std::uint64_t>(s) * kMaxRuns + r) * 3 + k] = base[k] + partial[r][k];
    unsigned long long steps = 0, interactions = 0;
#pragma unroll
    for (int r = 0; r < kMaxRuns; ++r) {
      steps += row[r * kShowerFields + 3];
      interactions += row[r * kShowerFields + 4];
#pragma unroll
      for (int k = 0; k < kShowerFields; ++k) row[r * kShowerFields + k] = 0;  // ready for the next front
    }
    f.history[s] = history + within[2];
    unsigned long long* st = f.stats + static_cast<std::uint64_t>(s) * kSlotStats;
    st[photon ? 0 : 1] += steps;
    st[photon ? 2 : 3] += interactions;
    st[4] += within[2];
    f.report[kReportHeader + 2 * static_cast<std::uint64_t>(s)] = within[0];
    f.report[kReportHeader + 2 * static_cast<std::uint64_t>(s) + 1] = within[1];
  }
  if (threadIdx.x == 0) {
    FrontTotals out{};
    out.next = carry[0];
    out.other = carry[1];
    out.children = carry[2];
    out.steps = carry[3];
    out.interactions = carry[4];
    *reinterpret_cast<FrontTotals*>(f.report) = out;
  }
}

__global__ __launch_bounds__(kStepBlock) void scatter_kernel(FrontArgs f, ParticleQueue next_queue,
                                                             ParticleQueue other_queue,
                                                             std::uint64_t other_base) {
  __shared__ std::uint32_t wave_sums[kStepBlock / 64][3];
  std::uint64_t const i = static_cast<std::uint64_t>(blockIdx.x) * kStepBlock + threadIdx.x;
  std::uint32_t const word = i < f.count ? f.packed[i] : 0u;
  std::uint32_t const v[3] = {word & 3u, (word >> 2) & 3u, (word >> 4) & 3u};
  std::uint32_t const wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  std::uint32_t inc[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) inc[k] = wave_inclusive(v[k], lane);
  if (lane == 63)
    for (int k = 0; k < 3; ++k) wave_sums[wave][k] = inc[k];
  __syncthreads();
  if (i >= f.count) return;
  std::uint32_t const s = word >> 12;
  unsigned long long const* adj = f.adjust + (static_cast<std::uint64_t>(s) * kMaxRuns + run_of(f, i)) * 3;
  std::uint64_t pos[3];
  for (int k = 0; k < 3; ++k) {
    std::uint32_t before = 0;
    for (std::uint32_t w = 0; w < wave; ++w) before += wave_sums[w][k];
    pos[k] = f.offsets[3 * static_cast<std::uint64_t>(blockIdx.x) + k] + before + inc[k] - v[k] + adj[k];
  }
  std::uint64_t next_pos = pos[0], other_pos = other_base + pos[1];
  std::uint64_t const child_base = pos[2];
  for (int k = 0; k < 2; ++k) {
    std::uint32_t const kind = (word >> (6 + 2 * k)) & 3u;
    if (kind == kSlotNone) continue;
    c8mi_particle_state p = f.staging[2 * i + k];
    if (kind != kSlotNextContinue) p.history_id = child_base + p.history_id;
    if (kind == kSlotOtherChild) {
      store_particle(other_queue, other_pos, p);
      other_queue.shower[other_pos++] = s;
    } else {
      store_particle(next_queue, next_pos, p);
      next_queue.shower[next_pos++] = s;
    }
  }
}

unsigned grid_for(std::uint64_t n, int block) {
  std::uint64_t const g = (n + block - 1) / block;
  if (g > 0x7fffffffULL) throw std::length_error("front exceeds the launch grid");
  return static_cast<unsigned>(g);
}

}  // namespace

std::size_t step_lds_bytes(std::uint32_t bins) {
  return 2 * (4u * static_cast<std::size_t>(bins) + kCounters) * sizeof(long long);
}

std::size_t block_count_bytes(std::uint64_t n) { return grid_for(n, kStepBlock) * sizeof(BlockCounts); }

std::size_t offset_bytes(std::uint64_t n) { return grid_for(n, kStepBlock) * 3 * sizeof(unsigned long long); }

void launch_step(bool photon, EmContext const* context, std::uint32_t bins, FrontArgs const& f,
                 hipStream_t stream) {
  if (f.count == 0) return;
  unsigned const grid = grid_for(f.count, kStepBlock);
  std::size_t const lds = step_lds_bytes(bins);
  if (photon) step_kernel<true><<<grid, kStepBlock, lds, stream>>>(context, f);
  else step_kernel<false><<<grid, kStepBlock, lds, stream>>>(context, f);
  C8MI_HIP(hipGetLastError());
}

void launch_scan(bool photon, FrontArgs const& f, hipStream_t stream) {
  if (f.slots > static_cast<std::uint32_t>(kScanBlock) || f.slots > kMaxShowerSlots)
    throw std::length_error("em: more pool slots than one scan workgroup covers");
  std::uint32_t const blocks = f.count == 0 ? 0u : grid_for(f.count, kStepBlock);
  scan_kernel<<<1, kScanBlock, 0, stream>>>(f, blocks, photon ? 1 : 0);
  C8MI_HIP(hipGetLastError());
}

void launch_scatter(FrontArgs const& f, ParticleQueue const& next_queue, ParticleQueue const& other_queue,
                    std::uint64_t other_base, hipStream_t stream) {
  if (f.count == 0) return;
  scatter_kernel<<<grid_for(f.count, kStepBlock), kStepBlock, 0, stream>>>(f, next_queue, other_queue,
                                                                           other_base);
  C8MI_HIP(hipGetLastError());
}

}  // namespace c8mi::em
