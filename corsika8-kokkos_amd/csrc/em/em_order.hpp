// Ordering of a front before its step kernels (em_order.hip): the
// reference's wavefront bucketing (bucket_wavefront,
// corsika/detail/kokkos/scheduling/WavefrontBucketing.inl:356-466, applied
// before every front at PhotonCascade.inl:318-332 and LeptonCascade.inl:296-307)
// extended to a pool of showers and to the order of a call's input.
//
// A stable two-way partition of the queue indices moves the parked entries
// (showers whose call does not run this front) behind the participants, in
// queue order; one stable LSD radix sort of (key, queue index) pairs over the
// participants follows, then a gather into the front buffer. The key, most
// significant first:
//   slot            the pool slot: every shower's particles are contiguous;
//   bucket          wavefront_bucket_key (PID bucket, medium, energy bin) for
//                   showers whose front holds at least 256 particles, else 0;
//   class (2 bits)  for showers opening a call: pending (appended by the other
//                   species' fronts) before remaining (left by the species'
//                   previous call) before host-staged, the order of the
//                   reference's call input (WavefrontQueue.inl:207-260).
// Ties keep the queue order. Field widths are chosen per front by the host;
// absent fields take no bits. The parked entries are carried over to the next
// queue unchanged; their order across showers does not matter (a front after a
// parked one is always ordered again) and within a shower it is kept.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "em/em_device.hpp"

namespace c8mi::em {

// Slot word of a queued particle: the pool slot in the low bits, the call-input
// class in the top two. A front reads only clean words (class remaining = 0).
inline constexpr std::uint32_t kSlotBits = 0xffffu;
inline constexpr int kClassShift = 30;
inline constexpr std::uint32_t kClassRemaining = 0;
inline constexpr std::uint32_t kClassPending = 1;
inline constexpr std::uint32_t kClassHost = 2;

// Fronts of a shower at or above this size are bucketed (MinimumWavefrontRadixSortSize,
// corsika/kokkos/scheduling/WavefrontBucketing.hpp:46).
inline constexpr std::uint64_t kBucketMinimum = 256;

// Per-slot control of one front, written by the host.
inline constexpr std::uint32_t kControlRun = 1u;      // the shower's call runs this front
inline constexpr std::uint32_t kControlBucket = 2u;   // sort its particles by bucket key
inline constexpr std::uint32_t kControlOpen = 4u;     // its call opens: order by class, set history
struct SlotControl {
  std::uint32_t flags;
  std::uint32_t pad;
  unsigned long long first_history;  // kControlOpen: the call's first secondary history ID
};

// Bit layout of one front's sort key.
struct OrderLayout {
  std::uint32_t class_bits;   // 2 when a shower opens a call, else 0
  std::uint32_t bucket_bits;  // 16 energy + medium rank bits + 2 PID, when a shower is bucketed
  std::uint32_t medium_bits;
  std::uint32_t slot_bits;
  std::uint32_t parked;       // 1 when a shower sits this front out: partition first
  std::uint32_t total() const { return class_bits + bucket_bits + slot_bits; }
};

class FrontOrder {
 public:
  FrontOrder() = default;
  ~FrontOrder();
  FrontOrder(FrontOrder const&) = delete;
  FrontOrder& operator=(FrontOrder const&) = delete;

  // Sorted distinct medium IDs (signed order) that particles may carry.
  void set_media(std::int32_t const* media, std::uint32_t count, hipStream_t stream);
  std::uint32_t medium_bits() const noexcept { return medium_bits_; }

  // Orders queue entries [0, n) of `in` into `out`: the `participants`
  // entries of running showers first (the host counts them from the same
  // control words), then the parked entries, which keep their slot words. Also
  // applies kControlOpen history IDs (slots < slot_count).
  void order(ParticleQueue const& in, ParticleQueue const& out, std::uint64_t n, std::uint64_t participants,
             SlotControl const* control, std::uint32_t slot_count, unsigned long long* history,
             OrderLayout const& layout, hipStream_t stream);

 private:
  void reserve(std::uint64_t n, std::uint32_t bits, bool partition, hipStream_t stream);

  std::int32_t* media_ = nullptr;
  std::uint32_t media_count_ = 0;
  std::uint32_t medium_bits_ = 0;
  unsigned long long* keys_[2] = {nullptr, nullptr};
  std::uint32_t* values_[2] = {nullptr, nullptr};
  std::uint8_t* run_flags_ = nullptr;           // partition flags, one per queue entry
  unsigned int* selected_count_ = nullptr;      // the partition's count (the host knows it)
  void* temporary_ = nullptr;
  std::size_t temporary_bytes_ = 0;
  std::uint64_t capacity_ = 0;
};

// Copies the parked tail [first, n) of `front` to `dest` starting at entry
// *dest_start, a device word (the front's count of continuing particles, which
// the scan kernel writes).
void launch_carry(ParticleQueue const& front, std::uint64_t first, std::uint64_t n, ParticleQueue const& dest,
                  unsigned long long const* dest_start, hipStream_t stream);

}  // namespace c8mi::em
