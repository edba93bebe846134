// Qualifiers of the step path: every function one particle step calls is
// compiled for the device kernels and for the host's scalar stepper (the CPU
// steps of the handoff prefix), from the same source, so both sides run the
// same arithmetic.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#define C8MI_HD __host__ __device__ __forceinline__
#define C8MI_HD_NOINLINE __host__ __device__ __noinline__ inline

namespace c8mi::em {

// The classification functions of both sides (the device overloads live in
// std beside the host ones).
using std::isfinite;
using std::isinf;
using std::isnan;

// Counter increment of the step path: one device atomic in a kernel; a plain
// add in the host stepper, whose accumulators belong to one thread.
C8MI_HD unsigned long long fetch_add(unsigned long long* address, unsigned long long value) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicAdd(address, value);
#else
  unsigned long long const old = *address;
  *address = old + value;
  return old;
#endif
}

}  // namespace c8mi::em
