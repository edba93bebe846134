// Front ordering: parked partition, sort keys, stable radix sort, gather;
// parked-tail carry. See em_order.hpp for the key layout.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_partition.hpp>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>

#include <cmath>
#include <cstdint>
#include <stdexcept>

#include "em/em_order.hpp"
#include "runtime.hpp"

namespace c8mi::em {

namespace {

constexpr int kOrderBlock = 256;

// rocPRIM's radix sort: fronts below this many participants take its merge
// sort, larger ones its onesweep radix passes. Lower limits (2^16, 2^13) were
// measured slower on every EM workload (10 TeV with radio: up to 10%).
constexpr std::size_t kMergeSortLimit = std::size_t{1} << 20;
using SortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                              rocprim::default_config, kMergeSortLimit>;

unsigned order_grid(std::uint64_t n) {
  std::uint64_t const g = (n + kOrderBlock - 1) / kOrderBlock;
  if (g > 0x7fffffffULL) throw std::length_error("em: front exceeds the ordering grid");
  return static_cast<unsigned>(g ? g : 1);
}

// wavefront_bucket_key's energy bin (WavefrontBucketing.inl:33-41):
// floor(log2(E / GeV) * 16 + 32768) clamped to [0, 65534]; 65535 for a
// non-positive or non-finite energy.
__device__ __forceinline__ std::uint32_t energy_bin(double e) {
  if (!(e > 0.) || !isfinite(e)) return 65535u;
  double const c = floor(log2(e) * 16. + 32768.);
  if (!(c > 0.)) return 0u;
  if (c >= 65534.) return 65534u;
  return static_cast<std::uint32_t>(c);
}

// pid_bucket (WavefrontBucketing.inl:24-30).
__device__ __forceinline__ std::uint32_t pid_bucket(std::int32_t pid) {
  return pid == 22 ? 0u : pid == 11 ? 1u : pid == -11 ? 2u : 3u;
}

// Rank of `medium` among the sorted distinct media (signed order, which is the
// order of medium ^ 2^31 in the reference's key).
__device__ __forceinline__ std::uint32_t medium_rank(std::int32_t medium, std::int32_t const* media,
                                                     std::uint32_t count) {
  std::uint32_t lo = 0, hi = count;
  while (lo < hi) {
    std::uint32_t const mid = (lo + hi) / 2;
    if (media[mid] < medium) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Call-input order of the classes: pending, remaining, host.
__device__ __forceinline__ std::uint32_t class_rank(std::uint32_t cls) {
  return cls == kClassPending ? 0u : cls == kClassRemaining ? 1u : 2u;
}

// Applies the kControlOpen history IDs and flags the entries of running
// showers for the partition.
__global__ __launch_bounds__(kOrderBlock) void order_flags_kernel(ParticleQueue in, std::uint64_t n,
                                                                  SlotControl const* __restrict__ control,
                                                                  std::uint32_t slot_count,
                                                                  unsigned long long* __restrict__ history,
                                                                  std::uint8_t* __restrict__ run) {
  std::uint64_t const i = static_cast<std::uint64_t>(blockIdx.x) * kOrderBlock + threadIdx.x;
  if (i < slot_count && (control[i].flags & kControlOpen)) history[i] = control[i].first_history;
  if (i >= n) return;
  run[i] = (control[in.shower[i] & kSlotBits].flags & kControlRun) ? 1u : 0u;
}

// Sort keys of `count` entries: entry j is queue entry selected[j] (the
// partition's participants, whose indices are already in `values`), or j
// itself when `selected` is null (then `values` gets j, and `history` the
// kControlOpen IDs).
__global__ __launch_bounds__(kOrderBlock) void order_keys_kernel(
    ParticleQueue in, std::uint64_t count, std::uint32_t const* __restrict__ selected,
    SlotControl const* __restrict__ control, std::uint32_t slot_count, unsigned long long* __restrict__ history,
    std::int32_t const* __restrict__ media, std::uint32_t media_count, OrderLayout layout, unsigned long long* __restrict__ keys,
    std::uint32_t* __restrict__ values) {
  std::uint64_t const j = static_cast<std::uint64_t>(blockIdx.x) * kOrderBlock + threadIdx.x;
  if (history && j < slot_count && (control[j].flags & kControlOpen)) history[j] = control[j].first_history;
  if (j >= count) return;
  std::uint64_t const i = selected ? selected[j] : j;
  std::uint32_t const word = in.shower[i];
  std::uint32_t const slot = word & kSlotBits;
  std::uint32_t const flags = control[slot].flags;
  bool const run = (flags & kControlRun) != 0;
  unsigned long long key = 0;
  std::uint32_t shift = 0;
  if (layout.class_bits) {
    if (run && (flags & kControlOpen)) key |= class_rank(word >> kClassShift);
    shift += layout.class_bits;
  }
  if (layout.bucket_bits) {
    if (run && (flags & kControlBucket)) {
      unsigned long long const bucket =
          (static_cast<unsigned long long>(pid_bucket(in.pid[i])) << (16 + layout.medium_bits)) |
          (static_cast<unsigned long long>(medium_rank(in.medium[i], media, media_count)) << 16) |
          energy_bin(in.energy[i]);
      key |= bucket << shift;
    }
    shift += layout.bucket_bits;
  }
  if (layout.slot_bits) key |= static_cast<unsigned long long>(slot) << shift;
  keys[j] = key;
  if (!selected) values[j] = static_cast<std::uint32_t>(i);
}

__global__ __launch_bounds__(kOrderBlock) void order_gather_kernel(ParticleQueue in, ParticleQueue out,
                                                                   std::uint32_t const* __restrict__ order,
                                                                   std::uint64_t n,
                                                                   SlotControl const* __restrict__ control) {
  std::uint64_t const j = static_cast<std::uint64_t>(blockIdx.x) * kOrderBlock + threadIdx.x;
  if (j >= n) return;
  std::uint64_t const src = order[j];
  std::uint32_t const word = in.shower[src];
  std::uint32_t const slot = word & kSlotBits;
  store_particle(out, j, load_particle(in, src));
  // a front reads clean slot words; parked particles keep their class
  out.shower[j] = (control[slot].flags & kControlRun) ? slot : word;
}

__global__ __launch_bounds__(kOrderBlock) void carry_kernel(ParticleQueue front, std::uint64_t first,
                                                            std::uint64_t n, ParticleQueue dest,
                                                            unsigned long long const* __restrict__ dest_start) {
  std::uint64_t const j = first + static_cast<std::uint64_t>(blockIdx.x) * kOrderBlock + threadIdx.x;
  if (j >= n) return;
  std::uint64_t const at = *dest_start + (j - first);
  store_particle(dest, at, load_particle(front, j));
  dest.shower[at] = front.shower[j];
}

}  // namespace

FrontOrder::~FrontOrder() {
  for (auto* p : {static_cast<void*>(media_), static_cast<void*>(keys_[0]), static_cast<void*>(keys_[1]),
                  static_cast<void*>(values_[0]), static_cast<void*>(values_[1]), static_cast<void*>(run_flags_),
                  static_cast<void*>(selected_count_), temporary_})
    if (p) (void)hipFree(p);
}

void FrontOrder::set_media(std::int32_t const* media, std::uint32_t count, hipStream_t stream) {
  if (count > media_count_ || media_ == nullptr) {
    C8MI_HIP(hipStreamSynchronize(stream));
    if (media_) C8MI_HIP(hipFree(media_));
    media_ = nullptr;
    C8MI_HIP(hipMalloc(reinterpret_cast<void**>(&media_), (count ? count : 1) * sizeof(std::int32_t)));
  }
  if (count)
    C8MI_HIP(hipMemcpyAsync(media_, media, count * sizeof(std::int32_t), hipMemcpyHostToDevice, stream));
  C8MI_HIP(hipStreamSynchronize(stream));  // `media` is the caller's host array
  media_count_ = count;
  medium_bits_ = 0;
  while ((1u << medium_bits_) < count) ++medium_bits_;
}

void FrontOrder::reserve(std::uint64_t n, std::uint32_t bits, bool partition, hipStream_t stream) {
  if (n > 0xffffffffULL) throw std::length_error("em: front of 2^32 or more particles");
  if (n > capacity_) {
    C8MI_HIP(hipStreamSynchronize(stream));  // earlier fronts may still read the arrays
    std::uint64_t cap = capacity_ ? capacity_ : 4096;
    while (cap < n) cap *= 2;
    for (int b = 0; b < 2; ++b) {
      if (keys_[b]) C8MI_HIP(hipFree(keys_[b]));
      if (values_[b]) C8MI_HIP(hipFree(values_[b]));
      keys_[b] = nullptr;
      values_[b] = nullptr;
      C8MI_HIP(hipMalloc(reinterpret_cast<void**>(&keys_[b]), cap * sizeof(unsigned long long)));
      C8MI_HIP(hipMalloc(reinterpret_cast<void**>(&values_[b]), cap * sizeof(std::uint32_t)));
    }
    if (run_flags_) C8MI_HIP(hipFree(run_flags_));
    run_flags_ = nullptr;
    C8MI_HIP(hipMalloc(reinterpret_cast<void**>(&run_flags_), cap));
    capacity_ = cap;
  }
  if (!selected_count_) C8MI_HIP(hipMalloc(reinterpret_cast<void**>(&selected_count_), sizeof(unsigned int)));
  std::size_t need = 0, part = 0;
  auto const items = static_cast<std::uint32_t>(capacity_);
  // 64-bit keys also for keys of 32 bits or fewer: measured no slower
  if (bits > 0)
    C8MI_HIP(rocprim::radix_sort_pairs<SortConfig>(nullptr, need, keys_[0], keys_[1], values_[0], values_[1], items, 0u, 64u,
                                       stream));
  if (partition)
    C8MI_HIP(rocprim::partition_two_way(nullptr, part, rocprim::counting_iterator<std::uint32_t>(0u), run_flags_,
                                        values_[0], values_[1], selected_count_, capacity_, stream));
  need = std::max(need, part);
  if (need > temporary_bytes_) {
    C8MI_HIP(hipStreamSynchronize(stream));
    if (temporary_) C8MI_HIP(hipFree(temporary_));
    temporary_ = nullptr;
    C8MI_HIP(hipMalloc(&temporary_, need));
    temporary_bytes_ = need;
  }
}

void FrontOrder::order(ParticleQueue const& in, ParticleQueue const& out, std::uint64_t n,
                       std::uint64_t participants, SlotControl const* control, std::uint32_t slot_count,
                       unsigned long long* history, OrderLayout const& layout, hipStream_t stream) {
  std::uint32_t const bits = layout.total();
  if (bits > 64) throw std::logic_error("em: front sort key wider than 64 bits");
  bool const partition = layout.parked != 0;
  if (participants > n || (!partition && participants != n))
    throw std::logic_error("em: front participants do not match the queued entries");
  reserve(n, bits, partition, stream);
  std::uint64_t const m = partition ? participants : n;
  auto keys = [&](std::uint64_t count, std::uint32_t const* selected, unsigned long long* ids, std::uint64_t threads) {
    if (threads == 0) return;
    order_keys_kernel<<<order_grid(threads), kOrderBlock, 0, stream>>>(in, count, selected, control, slot_count, ids,
                                                                         media_, media_count_, layout, keys_[0],
                                                                         values_[0]);
    C8MI_HIP(hipGetLastError());
  };
  if (partition) {
    // participants' indices to values_[0][0, m), the parked ones' (in queue
    // order) straight to their final place values_[1][m, n)
    order_flags_kernel<<<order_grid(std::max<std::uint64_t>(n, slot_count)), kOrderBlock, 0, stream>>>(
        in, n, control, slot_count, history, run_flags_);
    C8MI_HIP(hipGetLastError());
    if (n == 0) return;
    std::size_t bytes = temporary_bytes_;
    C8MI_HIP(rocprim::partition_two_way(temporary_, bytes, rocprim::counting_iterator<std::uint32_t>(0u), run_flags_,
                                        values_[0], values_[1] + m, selected_count_, n, stream));
    if (bits != 0) keys(m, values_[0], nullptr, m);
  } else {
    keys(n, nullptr, history, std::max<std::uint64_t>(n, slot_count));
    if (n == 0) return;
  }
  std::uint32_t const* order = values_[0];
  if (bits != 0 && m != 0) {
    std::size_t bytes = temporary_bytes_;
    auto const items = static_cast<std::uint32_t>(m);
    C8MI_HIP(rocprim::radix_sort_pairs<SortConfig>(temporary_, bytes, keys_[0], keys_[1], values_[0], values_[1], items, 0u, bits,
                                       stream));
    order = values_[1];
  } else if (partition) {
    if (m != 0)
      C8MI_HIP(hipMemcpyAsync(values_[1], values_[0], m * sizeof(std::uint32_t), hipMemcpyDeviceToDevice, stream));
    order = values_[1];
  }
  order_gather_kernel<<<order_grid(n), kOrderBlock, 0, stream>>>(in, out, order, n, control);
  C8MI_HIP(hipGetLastError());
}

void launch_carry(ParticleQueue const& front, std::uint64_t first, std::uint64_t n, ParticleQueue const& dest,
                  unsigned long long const* dest_start, hipStream_t stream) {
  if (n <= first) return;
  carry_kernel<<<order_grid(n - first), kOrderBlock, 0, stream>>>(front, first, n, dest, dest_start);
  C8MI_HIP(hipGetLastError());
}

}  // namespace c8mi::em
