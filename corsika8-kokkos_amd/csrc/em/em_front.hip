// Wavefront kernels of the EM backend. One front of one species is a short
// chain of launches on the backend's stream:
//   step       one thread per queued particle. Photons: the whole step
//              (selection, transport, vertex, final state) of PhotonFront.inl
//              fused into one kernel. e+-: the transport part (selection,
//              continuous step, transport, profile) of LeptonFront.inl; a step
//              that ends at an interaction vertex is compacted into a vertex
//              list (wave ballot, one atomic per workgroup), and
//   vertex     one thread per listed vertex runs the vertex reselection and the
//              final state, so the ~7% of e+- steps that interact run in full
//              wavefronts instead of holding every wavefront of the transport
//              kernel.
//              Up to two outputs per source go to a per-source staging pair
//              and a packed count word; profile and ledger contributions
//              accumulate in workgroup-private LDS copies (one for the shower
//              of the workgroup's first source, one for its last) that are
//              added to the showers' device totals once per workgroup.
//   count      per-workgroup output counts and per-(slot, run) output totals
//              from the packed words.
//   scan       one workgroup: exclusive scan of the per-workgroup counts, and
//              per slot the position and history offsets of every run's
//              outputs.
//   scatter    stable compaction: the block prefix of the per-source counts
//              plus the block offset gives the source's global output rank;
//              the (slot, run) offset turns it into its rank inside its own
//              shower, so each shower's queue segment, order and history IDs
//              (first + child rank + k) are those of the shower run alone, as
//              the reference's stable scans give them.
// Queues are sorted by slot within each run (the previous front's output is
// one run, the other species' secondaries and host-staged particles each
// append one).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "em/em_device.hpp"
#include "em/em_front.hpp"
#include "em/em_order.hpp"
#include "em/steps.hpp"
#include "runtime.hpp"

namespace c8mi::em {

namespace {

constexpr int kStepBlock = 256;
constexpr int kScanBlock = 1024;
constexpr std::uint32_t kNoKey = 0xffffffffu;
constexpr int kRunShift = 20;
// Words of a slot's adjust row beyond the run offsets: its first source in the
// front + 1 (0: none) and that source's in-workgroup exclusive children prefix.
constexpr int kAdjustFirst = 3;
constexpr int kAdjustChildPrefix = 4;
// Vertex kernel grid: one workgroup per 2048 front particles (a chunk of 256
// vertices per workgroup and pass), more vertices are covered by the stride loop.
constexpr std::uint64_t kSourcesPerVertexBlock = 2048;

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

// The context (tables, environment and physics parameters, Molière coefficients)
// is read by every lane at uniform addresses throughout a step; a workgroup
// copies it into the front of its LDS once, so those reads are LDS reads
// instead of per-lane global loads.
inline constexpr std::uint32_t kContextWords = (sizeof(EmContext) + 7) / 8;
static_assert(alignof(EmContext) <= 8, "context words");

__device__ __forceinline__ EmContext const* stage_context(EmContext const* __restrict__ context, long long* lds) {
  auto const* src = reinterpret_cast<long long const*>(context);
  for (std::uint32_t k = threadIdx.x; k < kContextWords; k += blockDim.x) lds[k] = src[k];
  return reinterpret_cast<EmContext const*>(lds);
}

// Workgroup-private LDS copies of the profile and ledger of up to two showers.
struct LdsShowers {
  long long* base;
  std::uint32_t stride;  // 4 * bins + kCounters
  std::uint32_t sh0, sh1;
  std::uint32_t copies;
};

__device__ __forceinline__ LdsShowers lds_showers(long long* lds, std::uint32_t bins, std::uint32_t sh0,
                                                  std::uint32_t sh1) {
  LdsShowers l{lds, 4u * bins + kCounters, sh0, sh1, sh0 == sh1 ? 1u : 2u};
  for (std::uint32_t k = threadIdx.x; k < l.copies * l.stride; k += blockDim.x) lds[k] = 0;
  return l;
}

__device__ __forceinline__ phys::Accumulator accumulator(EmContext const* context, LdsShowers const& l,
                                                         std::uint32_t s) {
  EmContext const& c = *context;
  std::uint32_t const bins = c.profile_bins;
  long long* hist = s == l.sh0 ? l.base : s == l.sh1 ? l.base + l.stride : nullptr;
  return phys::Accumulator{hist ? hist : c.profile + static_cast<std::uint64_t>(s) * 4u * bins,
                           hist ? hist + 4u * bins : c.counters + static_cast<std::uint64_t>(s) * kCounters,
                           context, &c.showers[s], s};
}

// Adds the workgroup's LDS copies to the showers' device totals (after a barrier).
__device__ __forceinline__ void flush(EmContext const& c, LdsShowers const& l) {
  std::uint32_t const bins = c.profile_bins;
  for (std::uint32_t copy = 0; copy < l.copies; ++copy) {
    std::uint32_t const s = copy ? l.sh1 : l.sh0;
    long long const* src = l.base + copy * l.stride;
    long long* hist = c.profile + static_cast<std::uint64_t>(s) * 4u * bins;
    long long* counters = c.counters + static_cast<std::uint64_t>(s) * kCounters;
    for (std::uint32_t k = threadIdx.x; k < l.stride; k += blockDim.x) {
      long long const value = src[k];
      if (value == 0) continue;
      long long* dst = k < 4u * bins ? hist + k : counters + (k - 4u * bins);
      atomicAdd(reinterpret_cast<unsigned long long*>(dst), static_cast<unsigned long long>(value));
    }
  }
}

__device__ __forceinline__ void write_outputs(FrontArgs const& f, std::uint64_t i, phys::Outputs const& o,
                                              std::uint32_t s) {
  f.packed[i] = pack_counts(o.next, o.other, o.children, o.kind[0], o.kind[1], o.step, o.interaction, s);
}

__device__ __forceinline__ phys::Outputs no_outputs(FrontArgs const& f, std::uint64_t i) {
  phys::Outputs o;
  o.staging = f.staging + 2 * i;
  o.track = f.tracks ? f.tracks + i : nullptr;
  o.track_stride = f.track_stride;
  o.kind[0] = o.kind[1] = kSlotNone;
  o.next = o.other = o.children = 0;
  o.step = o.interaction = o.tracked = false;
  return o;
}

__global__ __launch_bounds__(kStepBlock) void photon_step_kernel(EmContext const* __restrict__ context,
                                                                 FrontArgs f) {
  extern __shared__ long long lds[];
  EmContext const* const local = stage_context(context, lds);
  std::uint64_t const first = static_cast<std::uint64_t>(blockIdx.x) * kStepBlock;
  std::uint64_t const last = (first + kStepBlock < f.count ? first + kStepBlock : f.count) - 1;
  LdsShowers const l = lds_showers(lds + kContextWords, context->profile_bins, f.in.shower[first], f.in.shower[last]);
  __syncthreads();
  EmContext const& c = *local;
  std::uint64_t const i = first + threadIdx.x;
  if (i < f.count) {
    std::uint32_t const s = f.in.shower[i];
    phys::Accumulator const acc = accumulator(local, l, s);
    phys::Particle p;
    static_cast<c8mi_particle_state&>(p) = load_particle(f.in, i);
    p.shower_key = c.showers[s].key;
    phys::Outputs o = no_outputs(f, i);
    phys::photon_step(c, acc, p, o);
    write_outputs(f, i, o, s);
  }
  __syncthreads();
  flush(c, l);
}

// Three waves per SIMD (168 VGPRs, a little scratch) measured faster than two
// for the transport part; the photon and vertex kernels keep two and one.
__global__ __launch_bounds__(kStepBlock, 3) void lepton_transport_kernel(EmContext const* __restrict__ context,
                                                                      FrontArgs f) {
  extern __shared__ long long lds[];
  __shared__ std::uint32_t wave_vertices[kStepBlock / 64];
  __shared__ unsigned long long block_base;
  EmContext const* const local = stage_context(context, lds);
  std::uint64_t const first = static_cast<std::uint64_t>(blockIdx.x) * kStepBlock;
  std::uint64_t const last = (first + kStepBlock < f.count ? first + kStepBlock : f.count) - 1;
  LdsShowers const l = lds_showers(lds + kContextWords, context->profile_bins, f.in.shower[first], f.in.shower[last]);
  __syncthreads();
  EmContext const& c = *local;
  std::uint64_t const i = first + threadIdx.x;
  bool deferred = false;
  std::uint32_t s = 0;
  phys::LeptonVertex vx;
  if (i < f.count) {
    s = f.in.shower[i];
    phys::Accumulator const acc = accumulator(local, l, s);
    phys::Particle p;
    static_cast<c8mi_particle_state&>(p) = load_particle(f.in, i);
    p.shower_key = c.showers[s].key;
    phys::Outputs o = no_outputs(f, i);
    deferred = phys::lepton_transport(c, acc, p, o, vx);
    write_outputs(f, i, o, s);  // a deferred source's word is rewritten by the vertex kernel
    if (o.track && !o.step) phys::write_no_track(o.track, o.track_stride);
  }
  // compact the vertices of the workgroup, in source order
  std::uint32_t const wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  unsigned long long const mask = __ballot(deferred);
  if (lane == 0) wave_vertices[wave] = static_cast<std::uint32_t>(__popcll(mask));
  __syncthreads();
  if (threadIdx.x == 0) {
    std::uint32_t total = 0;
    for (int w = 0; w < kStepBlock / 64; ++w) total += wave_vertices[w];
    block_base = total ? atomicAdd(f.vertex_count, static_cast<unsigned long long>(total)) : 0ULL;
  }
  flush(c, l);
  __syncthreads();
  if (deferred) {
    std::uint32_t before = 0;
    for (std::uint32_t w = 0; w < wave; ++w) before += wave_vertices[w];
    unsigned long long const below = lane ? (mask & (~0ULL >> (64 - lane))) : 0ULL;
    VertexRecord r;
    r.parent = vx.parent;
    r.shower_key = vx.parent.shower_key;
    r.source = i;
    r.total_rate = vx.total_rate;
    r.process_uniform = vx.process_uniform;
    r.particle_mass_gev = vx.particle_mass_gev;
    r.vertex_density = vx.vertex_density;
    r.slot = s;
    r.pad[0] = r.pad[1] = r.pad[2] = 0;
    f.vertices[block_base + before + static_cast<std::uint32_t>(__popcll(below))] = r;
  }
}

__global__ __launch_bounds__(kStepBlock, 3) void lepton_move_kernel(EmContext const* __restrict__ context,
                                                                    FrontArgs f) {
  extern __shared__ long long lds[];
  EmContext const* const local = stage_context(context, lds);
  std::uint64_t const first = static_cast<std::uint64_t>(blockIdx.x) * kStepBlock;
  std::uint64_t const last = (first + kStepBlock < f.count ? first + kStepBlock : f.count) - 1;
  LdsShowers const l = lds_showers(lds + kContextWords, context->profile_bins, f.in.shower[first], f.in.shower[last]);
  __syncthreads();
  EmContext const& c = *local;
  std::uint64_t const i = first + threadIdx.x;
  if (i < f.count) {
    std::uint32_t const s = f.in.shower[i];
    phys::Accumulator const acc = accumulator(local, l, s);
    phys::Particle p;
    static_cast<c8mi_particle_state&>(p) = load_particle(f.in, i);
    p.shower_key = c.showers[s].key;
    phys::LeptonMove mv;
    if (i == 0) *f.moliere_open_count = 0;  // the previous front's tail kernel has finished
    bool const moved = phys::lepton_move(c, acc, p, mv);
    if (moved) static_cast<phys::LeptonMove*>(f.moves)[i] = mv;
    else f.packed[i] = pack_counts(0, 0, 0, kSlotNone, kSlotNone, false, false, s);
    f.moved[i] = moved ? kMovedOk : kMovedNone;
  }
  __syncthreads();
  flush(c, l);
}

// Molière scattering of the moved sources: the move record's end direction is
// deflected in place; a failed sampling marks the source (kMovedMoliereFailed)
// for the finish kernel; a sample not converged after the first Newton
// iterations is listed for the tail kernel. A kernel of its own: the sampler's
// registers do not weigh on the rest of the step. The context is read from
// global memory (its snapshot and table pointers are uniform, scalar loads).
template <bool Table>
__global__ __launch_bounds__(kStepBlock) void lepton_moliere_kernel(EmContext const* __restrict__ context,
                                                                    FrontArgs f) {
  std::uint64_t const i = static_cast<std::uint64_t>(blockIdx.x) * kStepBlock + threadIdx.x;
  if (i >= f.count || f.moved[i] != kMovedOk) return;
  auto* mv = static_cast<phys::LeptonMove*>(f.moves) + i;
  phys::MoliereOpen open;
  int const outcome = phys::moliere_scatter<Table>(*context, *mv, open);
  if (outcome == phys::kMoliereFailed) {
    f.moved[i] = kMovedMoliereFailed;
  } else if (outcome == phys::kMoliereOpen) {
    open.source = static_cast<std::uint32_t>(i);
    f.moved[i] = kMovedOpen;  // finished by the open-source kernels
    unsigned long long const k = atomicAdd(f.moliere_open_count, 1ULL);
    static_cast<phys::MoliereOpen*>(f.moliere_open)[k] = open;  // at most one per source
  }
}

// The open samples, one workgroup each (grid-stride over the list).
template <bool Table>
__global__ __launch_bounds__(moliere::kTailThreads) void lepton_moliere_tail_kernel(
    EmContext const* __restrict__ context, FrontArgs f) {
  __shared__ double jobs[moliere::kDeviceComponents * moliere::kJobSlots];
  unsigned long long const n = *f.moliere_open_count;
  for (unsigned long long w = blockIdx.x; w < n; w += gridDim.x) {
    phys::MoliereOpen const open = static_cast<phys::MoliereOpen const*>(f.moliere_open)[w];
    phys::LeptonMove mv = static_cast<phys::LeptonMove const*>(f.moves)[open.source];
    int const outcome = phys::moliere_scatter_wave<Table>(*context, mv, open, jobs);
    if (threadIdx.x == 0) {
      // the outcome goes with the open record: the source stays kMovedOpen for
      // the finish kernel of the other sources, which may run meanwhile
      auto* record = static_cast<phys::MoliereOpen*>(f.moliere_open) + w;
      if (outcome == phys::kMoliereFailed) {
        record->pad = kMovedMoliereFailed;
      } else {
        auto* out = static_cast<phys::LeptonMove*>(f.moves) + open.source;
        for (int a = 0; a < 3; ++a) out->end.direction[a] = mv.end.direction[a];
        record->pad = kMovedOk;
      }
    }
  }
}

// The finish of one moved source (Molière applied): outputs, profile, records
// and, for an interaction candidate, the vertex to list (returns true).
__device__ __forceinline__ bool finish_source(EmContext const* local, LdsShowers const& l, FrontArgs const& f,
                                              std::uint64_t i, std::uint32_t moved, std::uint32_t& s,
                                              phys::LeptonVertex& vx) {
  EmContext const& c = *local;
  s = f.in.shower[i];
  phys::Accumulator const acc = accumulator(local, l, s);
  phys::LeptonMove const mv = static_cast<phys::LeptonMove const*>(f.moves)[i];
  phys::Outputs o = no_outputs(f, i);
  auto const input = [&] {
    phys::Particle p;
    static_cast<c8mi_particle_state&>(p) = load_particle(f.in, i);
    p.shower_key = mv.start.shower_key;
    return p;
  };
  bool const deferred = phys::lepton_finish(c, acc, input, mv, moved == kMovedMoliereFailed, o, vx);
  write_outputs(f, i, o, s);
  if (f.tracks && !o.step) phys::write_no_track(f.tracks + i, f.track_stride);
  return deferred;
}

__device__ __forceinline__ VertexRecord vertex_record(phys::LeptonVertex const& vx, std::uint64_t i,
                                                      std::uint32_t s) {
  VertexRecord r;
  r.parent = vx.parent;
  r.shower_key = vx.parent.shower_key;
  r.source = i;
  r.total_rate = vx.total_rate;
  r.process_uniform = vx.process_uniform;
  r.particle_mass_gev = vx.particle_mass_gev;
  r.vertex_density = vx.vertex_density;
  r.slot = s;
  r.pad[0] = r.pad[1] = r.pad[2] = 0;
  return r;
}

// Two waves per SIMD (a little scratch) measured faster than one for the finish.
// Sources whose Molière sample is still open are left to lepton_finish_open_kernel.
__global__ __launch_bounds__(kStepBlock, 2) void lepton_finish_kernel(EmContext const* __restrict__ context,
                                                                   FrontArgs f) {
  extern __shared__ long long lds[];
  __shared__ std::uint32_t wave_vertices[kStepBlock / 64];
  __shared__ unsigned long long block_base;
  EmContext const* const local = stage_context(context, lds);
  std::uint64_t const first = static_cast<std::uint64_t>(blockIdx.x) * kStepBlock;
  std::uint64_t const last = (first + kStepBlock < f.count ? first + kStepBlock : f.count) - 1;
  LdsShowers const l = lds_showers(lds + kContextWords, context->profile_bins, f.in.shower[first], f.in.shower[last]);
  __syncthreads();
  EmContext const& c = *local;
  std::uint64_t const i = first + threadIdx.x;
  bool deferred = false;
  std::uint32_t s = 0;
  phys::LeptonVertex vx;
  std::uint32_t const moved = i < f.count ? f.moved[i] : kMovedOpen;
  if (moved == kMovedOk || moved == kMovedMoliereFailed) deferred = finish_source(local, l, f, i, moved, s, vx);
  else if (moved == kMovedNone && f.tracks) phys::write_no_track(f.tracks + i, f.track_stride);
  std::uint32_t const wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  unsigned long long const mask = __ballot(deferred);
  if (lane == 0) wave_vertices[wave] = static_cast<std::uint32_t>(__popcll(mask));
  __syncthreads();
  if (threadIdx.x == 0) {
    std::uint32_t total = 0;
    for (int w = 0; w < kStepBlock / 64; ++w) total += wave_vertices[w];
    block_base = total ? atomicAdd(f.vertex_count, static_cast<unsigned long long>(total)) : 0ULL;
  }
  flush(c, l);
  __syncthreads();
  if (deferred) {
    std::uint32_t before = 0;
    for (std::uint32_t w = 0; w < wave; ++w) before += wave_vertices[w];
    unsigned long long const below = lane ? (mask & (~0ULL >> (64 - lane))) : 0ULL;
    f.vertices[block_base + before + static_cast<std::uint32_t>(__popcll(below))] = vertex_record(vx, i, s);
  }
}

// The finish of the sources whose Molière sample went through the tail kernel
// (few, scattered over the front): one thread each, profile and counters
// straight to the showers' device totals, vertices into their own list.
__global__ __launch_bounds__(kStepBlock) void lepton_finish_open_kernel(EmContext const* __restrict__ context,
                                                                        FrontArgs f) {
  LdsShowers const l{nullptr, 0u, ~0u, ~0u, 0u};  // no LDS copies: every slot adds to its device totals
  unsigned long long const n = *f.moliere_open_count;
  for (unsigned long long w = static_cast<unsigned long long>(blockIdx.x) * kStepBlock + threadIdx.x; w < n;
       w += static_cast<unsigned long long>(gridDim.x) * kStepBlock) {
    phys::MoliereOpen const open = static_cast<phys::MoliereOpen const*>(f.moliere_open)[w];
    std::uint64_t const i = open.source;
    std::uint32_t s = 0;
    phys::LeptonVertex vx;
    if (finish_source(context, l, f, i, open.pad, s, vx))
      f.open_vertices[atomicAdd(f.open_vertex_count, 1ULL)] = vertex_record(vx, i, s);
  }
}

__global__ __launch_bounds__(kStepBlock) void lepton_vertex_kernel(EmContext const* __restrict__ context,
                                                                   FrontArgs f) {
  extern __shared__ long long lds[];
  EmContext const* const local = stage_context(context, lds);
  EmContext const& c = *local;
  std::uint64_t const vertices = *f.vertex_count;
  for (std::uint64_t first = static_cast<std::uint64_t>(blockIdx.x) * kStepBlock; first < vertices;
       first += static_cast<std::uint64_t>(gridDim.x) * kStepBlock) {
    std::uint64_t const last = (first + kStepBlock < vertices ? first + kStepBlock : vertices) - 1;
    LdsShowers const l =
        lds_showers(lds + kContextWords, context->profile_bins, f.vertices[first].slot, f.vertices[last].slot);
    __syncthreads();
    std::uint64_t const j = first + threadIdx.x;
    if (j < vertices) {
      VertexRecord const& r = f.vertices[j];
      std::uint32_t const s = r.slot;
      phys::Accumulator const acc = accumulator(local, l, s);
      phys::LeptonVertex vx;
      static_cast<c8mi_particle_state&>(vx.parent) = r.parent;
      vx.parent.shower_key = r.shower_key;
      vx.total_rate = r.total_rate;
      vx.process_uniform = r.process_uniform;
      vx.particle_mass_gev = r.particle_mass_gev;
      vx.vertex_density = r.vertex_density;
      phys::Outputs o = no_outputs(f, r.source);
      o.step = o.interaction = true;
      phys::lepton_vertex(c, acc, vx, o);
      write_outputs(f, r.source, o, s);
    }
    __syncthreads();
    flush(c, l);
    __syncthreads();  // the LDS copies are cleared for the next chunk
  }
}

// Per-workgroup totals and per-(slot, run) totals of the packed counts. Within
// a workgroup the keys are non-decreasing segments; the first source of a
// segment subtracts its exclusive block prefix and the last adds its
// inclusive prefix (arithmetic modulo 2^64).
__global__ __launch_bounds__(kStepBlock) void count_kernel(FrontArgs f) {
  __shared__ std::uint32_t keys[kStepBlock];
  __shared__ std::uint32_t wave_sums[kStepBlock / 64][kShowerFields];
  std::uint64_t const i = static_cast<std::uint64_t>(blockIdx.x) * kStepBlock + threadIdx.x;
  std::uint32_t const word = i < f.count ? f.packed[i] : 0u;
  std::uint32_t const v[kShowerFields] = {word & 3u, (word >> 2) & 3u, (word >> 4) & 3u, (word >> 10) & 1u,
                                          (word >> 11) & 1u};
  std::uint32_t const key = i < f.count ? (run_of(f, i) << kRunShift) | (word >> kPackedSlotShift) : kNoKey;
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
    if (blockIdx.x == 0) {  // the vertex kernels of this front have finished
      *f.vertex_count = 0;
      if (f.open_vertex_count) *f.open_vertex_count = 0;
    }
  }
  if (key == kNoKey) return;
  std::uint32_t const s = key & ((1u << kRunShift) - 1u), r = key >> kRunShift;
  unsigned long long* t = f.shower_totals + (static_cast<std::uint64_t>(s) * kMaxRuns + r) * kShowerFields;
  bool const start = threadIdx.x == 0 || keys[threadIdx.x - 1] != key;
  bool const end = threadIdx.x == kStepBlock - 1 || keys[threadIdx.x + 1] != key;
  if (!start && !end) return;
#pragma unroll
  for (int k = 0; k < kShowerFields; ++k) {
    std::uint32_t before = 0;
    for (std::uint32_t w = 0; w < wave; ++w) before += wave_sums[w][k];
    std::uint32_t const incl = before + inc[k], excl = incl - v[k];
    if (start && excl != 0) atomicAdd(&t[k], 0ULL - excl);
    if (end && incl != 0) atomicAdd(&t[k], static_cast<unsigned long long>(incl));
    // the shower's first source in the front: its in-workgroup exclusive
    // children prefix, from which the scan forms the shower's child offset
    if (k == 2 && start && (i == 0 || f.in.shower[i - 1] != s)) {
      unsigned long long* a = f.adjust + static_cast<std::uint64_t>(s) * kMaxRuns * 3;
      a[kAdjustFirst] = i + 1;
      a[kAdjustChildPrefix] = excl;
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

// Per slot (a front is one run sorted by slot, so a global output rank is
// already the output's queue position): the offset that turns a child's global
// rank into its history ID inside its shower (the shower's first children
// rank is its first source's workgroup offset plus the in-workgroup prefix the
// count kernel kept), new queue counts, history counters and statistics. One
// thread per slot.
__global__ __launch_bounds__(kStepBlock) void slot_update_kernel(FrontArgs f, int photon) {
  std::uint32_t const s = blockIdx.x * kStepBlock + threadIdx.x;
  if (s >= f.slots) return;
  unsigned long long* row = f.shower_totals + static_cast<std::uint64_t>(s) * kMaxRuns * kShowerFields;
  unsigned long long* adj = f.adjust + static_cast<std::uint64_t>(s) * kMaxRuns * 3;
  unsigned long long const next = row[0], other = row[1], children = row[2], steps = row[3],
                           interactions = row[4];
#pragma unroll
  for (int k = 0; k < kShowerFields; ++k) row[k] = 0;  // ready for the next front
  unsigned long long const history = f.history[s];
  unsigned long long const first = adj[kAdjustFirst];
  if (first != 0) {
    std::uint64_t const block = (first - 1) / kStepBlock;
    unsigned long long const excl = f.offsets[3 * block + 2] + adj[kAdjustChildPrefix];
    adj[0] = 0;
    adj[1] = 0;
    adj[2] = history - excl;
    adj[kAdjustFirst] = 0;
  }
  f.history[s] = history + children;
  unsigned long long* st = f.stats + static_cast<std::uint64_t>(s) * kSlotStats;
  st[photon ? 0 : 1] += steps;
  st[photon ? 2 : 3] += interactions;
  st[4] += children;
  f.report[kReportHeader + kReportPerSlot * static_cast<std::uint64_t>(s)] = next;
  f.report[kReportHeader + kReportPerSlot * static_cast<std::uint64_t>(s) + 1] = other;
  f.report[kReportHeader + kReportPerSlot * static_cast<std::uint64_t>(s) + 2] = children;
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
  std::uint32_t const s = word >> kPackedSlotShift;
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
    if (kind != kSlotNextContinue) {
      p.history_id = child_base + p.history_id;
      // children of generation-zero parents only (a handful per shower)
      if (p.generation == 1 && f.first_interactions)
        phys::patch_first_interaction(f.first_interactions[s], f.first_candidates[s], p);
    }
    if (kind == kSlotOtherChild) {
      store_particle(other_queue, other_pos, p);
      other_queue.shower[other_pos++] = s | (kClassPending << kClassShift);
    } else {
      store_particle(next_queue, next_pos, p);
      next_queue.shower[next_pos++] = s;
    }
  }
}

// Per-shower radio: the (slot, run) segments of a lepton front's sources. A
// run holds every slot at most once, contiguous, so each (slot, run) key has
// one first and one last source. The arrays are zero between uses; `end` > 0
// marks a present segment.
__global__ __launch_bounds__(kStepBlock) void radio_segments_kernel(FrontArgs f, unsigned long long* begin,
                                                                     unsigned long long* end) {
  std::uint64_t const i = static_cast<std::uint64_t>(blockIdx.x) * kStepBlock + threadIdx.x;
  if (i >= f.count) return;
  std::uint32_t const s = f.in.shower[i], r = run_of(f, i);
  std::uint64_t const key = static_cast<std::uint64_t>(s) * kMaxRuns + r;
  if (i == 0 || f.in.shower[i - 1] != s || run_of(f, i - 1) != r) begin[key] = i;
  if (i + 1 == f.count || f.in.shower[i + 1] != s || run_of(f, i + 1) != r) end[key] = i + 1;
}

// The segment list of the projection (one workgroup): every present segment
// of a slot with observers, in key order, with the offset of its first item
// (chunks x the slot's observers x window tiles); clears the key arrays.
__global__ __launch_bounds__(kScanBlock) void radio_items_kernel(unsigned long long* begin, unsigned long long* end,
                                                                  std::uint32_t keys,
                                                                  std::uint32_t const* observer_counts,
                                                                  std::uint64_t chunk, std::uint32_t tiles,
                                                                  radio::SegmentDev* segments,
                                                                  std::uint32_t* segment_count,
                                                                  unsigned long long* item_count) {
  __shared__ unsigned long long wave_sums[kScanBlock / 64][2];
  __shared__ unsigned long long carry[2];
  if (threadIdx.x < 2) carry[threadIdx.x] = 0;
  __syncthreads();
  std::uint32_t const wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  for (std::uint32_t base = 0; base < keys; base += kScanBlock) {
    std::uint32_t const e = base + threadIdx.x;
    unsigned long long b = 0, t = 0, items = 0;
    std::uint32_t slot = 0;
    if (e < keys) {
      t = end[e];
      if (t) {
        b = begin[e];
        slot = e / kMaxRuns;
        items = (t - b + chunk - 1) / chunk * observer_counts[slot] * tiles;
        begin[e] = end[e] = 0;
      }
    }
    unsigned long long const v[2] = {items, items ? 1ULL : 0ULL};
    unsigned long long inc[2];
    for (int k = 0; k < 2; ++k) inc[k] = wave_inclusive(v[k], lane);
    if (lane == 63)
      for (int k = 0; k < 2; ++k) wave_sums[wave][k] = inc[k];
    __syncthreads();
    unsigned long long before[2] = {0, 0}, all[2] = {0, 0};
    for (std::uint32_t w = 0; w < kScanBlock / 64; ++w)
      for (int k = 0; k < 2; ++k) {
        before[k] += w < wave ? wave_sums[w][k] : 0ULL;
        all[k] += wave_sums[w][k];
      }
    if (items) {
      radio::SegmentDev d;
      d.begin = b;
      d.end = t;
      d.item_begin = carry[0] + before[0] + inc[0] - v[0];
      d.slot = slot;
      d.pad = 0;
      segments[carry[1] + before[1] + inc[1] - v[1]] = d;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      carry[0] += all[0];
      carry[1] += all[1];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *item_count = carry[0];
    *segment_count = static_cast<std::uint32_t>(carry[1]);
  }
}

__global__ __launch_bounds__(256) void unstage_kernel(StagedParticle const* __restrict__ staged, std::uint64_t n,
                                                      ParticleQueue queue, std::uint64_t at) {
  std::uint64_t const i = static_cast<std::uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  store_particle(queue, at + i, staged[i].particle);
  queue.shower[at + i] = staged[i].slot | (kClassHost << kClassShift);
}

__global__ __launch_bounds__(256) void pack_kernel(ParticleQueue queue, std::uint64_t n,
                                                   c8mi_particle_state* __restrict__ out) {
  std::uint64_t const i = static_cast<std::uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) out[i] = load_particle(queue, i);
}

__global__ __launch_bounds__(256) void slot_init_kernel(PoolArrays pool, SlotInit const* __restrict__ inits) {
  SlotInit const init = inits[blockIdx.x];
  std::uint64_t const slot = init.slot;
  long long* hist = pool.profile + slot * 4u * pool.bins;
  for (std::uint32_t k = threadIdx.x; k < 4u * pool.bins; k += blockDim.x) hist[k] = 0;
  if (threadIdx.x < kCounters) pool.counters[slot * kCounters + threadIdx.x] = 0;
  if (threadIdx.x < kSlotStats) pool.stats[slot * kSlotStats + threadIdx.x] = 0;
  if (threadIdx.x < 2) pool.record_counts[2 * slot + threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    pool.first_candidates[slot] = 0;
    pool.history[slot] = init.first_history;
    pool.showers[slot] = init.shower;
  }
}

__global__ __launch_bounds__(128) void slot_gather_kernel(PoolArrays pool, std::uint32_t const* __restrict__ slots,
                                                          unsigned long long* __restrict__ out) {
  std::uint64_t const slot = slots[blockIdx.x];
  unsigned long long* o = out + static_cast<std::uint64_t>(blockIdx.x) * kGatherWords;
  std::uint32_t const k = threadIdx.x;
  if (k < kCounters) o[kGatherCounters + k] = static_cast<unsigned long long>(pool.counters[slot * kCounters + k]);
  else if (k < kCounters + kSlotStats) o[k] = pool.stats[slot * kSlotStats + (k - kCounters)];
  else if (k == kGatherHistory) o[k] = pool.history[slot];
  else if (k < kGatherRecords + 2) o[k] = pool.record_counts[2 * slot + (k - kGatherRecords)];
  else if (k < kGatherWords) o[k] = 0;
}

unsigned grid_for(std::uint64_t n, int block) {
  std::uint64_t const g = (n + block - 1) / block;
  if (g > 0x7fffffffULL) throw std::length_error("front exceeds the launch grid");
  return static_cast<unsigned>(g);
}

}  // namespace

std::size_t step_lds_bytes(std::uint32_t bins) {
  return (kContextWords + 2 * (4u * static_cast<std::size_t>(bins) + kCounters)) * sizeof(long long);
}

std::size_t block_count_bytes(std::uint64_t n) { return grid_for(n, kStepBlock) * sizeof(BlockCounts); }

std::size_t offset_bytes(std::uint64_t n) { return grid_for(n, kStepBlock) * 3 * sizeof(unsigned long long); }

void launch_photon_step(EmContext const* context, std::uint32_t bins, FrontArgs const& f, hipStream_t stream) {
  if (f.count == 0) return;
  photon_step_kernel<<<grid_for(f.count, kStepBlock), kStepBlock, step_lds_bytes(bins), stream>>>(context, f);
  C8MI_HIP(hipGetLastError());
}

void launch_lepton_transport(EmContext const* context, std::uint32_t bins, FrontArgs const& f, hipStream_t stream) {
  if (f.count == 0) return;
  lepton_transport_kernel<<<grid_for(f.count, kStepBlock), kStepBlock, step_lds_bytes(bins), stream>>>(context, f);
  C8MI_HIP(hipGetLastError());
}

std::size_t lepton_move_bytes() { return sizeof(phys::LeptonMove); }
std::size_t moliere_open_bytes() { return sizeof(phys::MoliereOpen); }

void launch_lepton_move(EmContext const* context, std::uint32_t bins, FrontArgs const& f, hipStream_t stream) {
  if (f.count == 0) return;
  lepton_move_kernel<<<grid_for(f.count, kStepBlock), kStepBlock, step_lds_bytes(bins), stream>>>(context, f);
  C8MI_HIP(hipGetLastError());
}

void launch_lepton_moliere(EmContext const* context, bool table, FrontArgs const& f, hipStream_t stream) {
  if (f.count == 0) return;
  if (table)
    lepton_moliere_kernel<true><<<grid_for(f.count, kStepBlock), kStepBlock, 0, stream>>>(context, f);
  else
    lepton_moliere_kernel<false><<<grid_for(f.count, kStepBlock), kStepBlock, 0, stream>>>(context, f);
  C8MI_HIP(hipGetLastError());
}

void launch_lepton_open(EmContext const* context, std::uint32_t bins, bool table, FrontArgs const& f,
                        hipStream_t stream) {
  if (f.count == 0) return;
  // the open samples are rare (~1e-5 of the steps): fixed grids of small
  // workgroups that leave at once when there are none
  constexpr unsigned kTailGrid = 256, kOpenGrid = 4;
  if (table)
    lepton_moliere_tail_kernel<true><<<kTailGrid, moliere::kTailThreads, 0, stream>>>(context, f);
  else
    lepton_moliere_tail_kernel<false><<<kTailGrid, moliere::kTailThreads, 0, stream>>>(context, f);
  lepton_finish_open_kernel<<<kOpenGrid, kStepBlock, 0, stream>>>(context, f);
  FrontArgs g = f;
  g.vertices = f.open_vertices;
  g.vertex_count = f.open_vertex_count;
  lepton_vertex_kernel<<<1, kStepBlock, step_lds_bytes(bins), stream>>>(context, g);
  C8MI_HIP(hipGetLastError());
}

void launch_radio_segments(FrontArgs const& f, std::uint32_t slots, unsigned long long* begin,
                           unsigned long long* end, std::uint32_t const* observer_counts, std::uint64_t chunk,
                           std::uint32_t tiles, radio::SegmentDev* segments, std::uint32_t* segment_count,
                           unsigned long long* item_count, hipStream_t stream) {
  if (f.count)
    radio_segments_kernel<<<grid_for(f.count, kStepBlock), kStepBlock, 0, stream>>>(f, begin, end);
  radio_items_kernel<<<1, kScanBlock, 0, stream>>>(begin, end, slots * kMaxRuns, observer_counts, chunk, tiles,
                                                   segments, segment_count, item_count);
  C8MI_HIP(hipGetLastError());
}

void launch_lepton_finish(EmContext const* context, std::uint32_t bins, FrontArgs const& f, hipStream_t stream) {
  if (f.count == 0) return;
  lepton_finish_kernel<<<grid_for(f.count, kStepBlock), kStepBlock, step_lds_bytes(bins), stream>>>(context, f);
  C8MI_HIP(hipGetLastError());
}

void launch_lepton_vertex(EmContext const* context, std::uint32_t bins, FrontArgs const& f, hipStream_t stream) {
  if (f.count == 0) return;
  unsigned const grid = grid_for((f.count + kSourcesPerVertexBlock - 1) / kSourcesPerVertexBlock, 1);
  lepton_vertex_kernel<<<grid, kStepBlock, step_lds_bytes(bins), stream>>>(context, f);
  C8MI_HIP(hipGetLastError());
}

void launch_count(FrontArgs const& f, hipStream_t stream) {
  if (f.count == 0) return;
  count_kernel<<<grid_for(f.count, kStepBlock), kStepBlock, 0, stream>>>(f);
  C8MI_HIP(hipGetLastError());
}

void launch_scan(bool photon, FrontArgs const& f, hipStream_t stream) {
  if (f.slots > kMaxShowerSlots) throw std::length_error("em: more pool slots than the scan covers");
  if (f.runs != 1) throw std::logic_error("em: a front's input must be one run sorted by slot");
  std::uint32_t const blocks = f.count == 0 ? 0u : grid_for(f.count, kStepBlock);
  scan_kernel<<<1, kScanBlock, 0, stream>>>(f, blocks, photon ? 1 : 0);
  C8MI_HIP(hipGetLastError());
  if (f.slots) {
    slot_update_kernel<<<grid_for(f.slots, kStepBlock), kStepBlock, 0, stream>>>(f, photon ? 1 : 0);
    C8MI_HIP(hipGetLastError());
  }
}

void launch_scatter(FrontArgs const& f, ParticleQueue const& next_queue, ParticleQueue const& other_queue,
                    std::uint64_t other_base, hipStream_t stream) {
  if (f.count == 0) return;
  scatter_kernel<<<grid_for(f.count, kStepBlock), kStepBlock, 0, stream>>>(f, next_queue, other_queue,
                                                                           other_base);
  C8MI_HIP(hipGetLastError());
}

void launch_unstage(StagedParticle const* staged, std::uint64_t n, ParticleQueue const& queue, std::uint64_t at,
                    hipStream_t stream) {
  if (n == 0) return;
  unstage_kernel<<<grid_for(n, 256), 256, 0, stream>>>(staged, n, queue, at);
  C8MI_HIP(hipGetLastError());
}

void launch_pack(ParticleQueue const& queue, std::uint64_t n, c8mi_particle_state* out, hipStream_t stream) {
  if (n == 0) return;
  pack_kernel<<<grid_for(n, 256), 256, 0, stream>>>(queue, n, out);
  C8MI_HIP(hipGetLastError());
}

void launch_slot_init(PoolArrays const& pool, SlotInit const* inits, std::uint32_t count, hipStream_t stream) {
  if (count == 0) return;
  slot_init_kernel<<<count, 256, 0, stream>>>(pool, inits);
  C8MI_HIP(hipGetLastError());
}

void launch_slot_gather(PoolArrays const& pool, std::uint32_t const* slots, std::uint32_t count,
                        unsigned long long* out, hipStream_t stream) {
  if (count == 0) return;
  static_assert(kGatherRecords + 2 <= kGatherWords && kGatherWords <= 128, "gather layout");
  slot_gather_kernel<<<count, 128, 0, stream>>>(pool, slots, out);
  C8MI_HIP(hipGetLastError());
}

}  // namespace c8mi::em
