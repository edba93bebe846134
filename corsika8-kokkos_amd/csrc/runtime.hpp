// c8mi runtime helpers: HIP error checking, per-thread last-error text for the
// C ABI, and small device-memory owners. Host C++ only (no kernels here).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

namespace c8mi {

// Error taxonomy of the reference's IEmBackend / RadioAccumulator
// (corsika/kokkos/em/execution/IEmBackend.hpp, BackendInstanceImplementation.inl:
// logic_error for state misuse, length_error for batches above a limit,
// runtime_error for device/budget/overflow failures, invalid_argument for
// configuration). The C ABI maps each to a distinct status code.
enum Status : int {
  kOk = 0,
  kInvalidArgument = 1,
  kLogicError = 2,
  kLengthError = 3,
  kRuntimeError = 4,
  kDeviceError = 5,
  kOverflowError = 6,
};

class DeviceError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

class OverflowError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

inline void check_hip(hipError_t status, char const* what) {
  if (status != hipSuccess)
    throw DeviceError(std::string(what) + ": " + hipGetErrorString(status));
}

#define C8MI_HIP(call) ::c8mi::check_hip((call), #call)

void set_last_error(std::string message);
char const* last_error();

// Owning device buffer (hipMalloc'd). Move-only.
template <class T>
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(std::size_t count) { allocate(count); }
  ~DeviceBuffer() { release(); }
  DeviceBuffer(DeviceBuffer const&) = delete;
  DeviceBuffer& operator=(DeviceBuffer const&) = delete;
  DeviceBuffer(DeviceBuffer&& other) noexcept
      : data_(std::exchange(other.data_, nullptr)), count_(std::exchange(other.count_, 0)) {}
  DeviceBuffer& operator=(DeviceBuffer&& other) noexcept {
    if (this != &other) {
      release();
      data_ = std::exchange(other.data_, nullptr);
      count_ = std::exchange(other.count_, 0);
    }
    return *this;
  }
  void allocate(std::size_t count) {
    release();
    if (count == 0) return;
    void* pointer = nullptr;
    C8MI_HIP(hipMalloc(&pointer, count * sizeof(T)));
    data_ = static_cast<T*>(pointer);
    count_ = count;
  }
  void release() noexcept {
    if (data_) (void)hipFree(data_);
    data_ = nullptr;
    count_ = 0;
  }
  T* data() const noexcept { return data_; }
  std::size_t size() const noexcept { return count_; }
  std::size_t bytes() const noexcept { return count_ * sizeof(T); }

 private:
  T* data_ = nullptr;
  std::size_t count_ = 0;
};

// Owning pinned host buffer (hipHostMalloc'd). Move-only.
template <class T>
class PinnedBuffer {
 public:
  PinnedBuffer() = default;
  explicit PinnedBuffer(std::size_t count) {
    if (count == 0) return;
    void* pointer = nullptr;
    C8MI_HIP(hipHostMalloc(&pointer, count * sizeof(T), hipHostMallocDefault));
    data_ = static_cast<T*>(pointer);
    count_ = count;
  }
  ~PinnedBuffer() {
    if (data_) (void)hipHostFree(data_);
  }
  PinnedBuffer(PinnedBuffer const&) = delete;
  PinnedBuffer& operator=(PinnedBuffer const&) = delete;
  PinnedBuffer(PinnedBuffer&& other) noexcept
      : data_(std::exchange(other.data_, nullptr)), count_(std::exchange(other.count_, 0)) {}
  PinnedBuffer& operator=(PinnedBuffer&& other) noexcept {
    if (this != &other) {
      if (data_) (void)hipHostFree(data_);
      data_ = std::exchange(other.data_, nullptr);
      count_ = std::exchange(other.count_, 0);
    }
    return *this;
  }
  T* data() const noexcept { return data_; }
  std::size_t size() const noexcept { return count_; }

 private:
  T* data_ = nullptr;
  std::size_t count_ = 0;
};

}  // namespace c8mi
