// Device-resident copy of a format-v7 native table set (the tables the
// reference's BackendInstance::initialize copies from a ProposalNativeTableSet,
// corsika/kokkos/em/tables/ProposalNativeTable.hpp:21-56). One allocation holds
// every array; view() returns the c8mi_native_table_set of device pointers that
// kernels take by value.
#pragma once

#include <cstdint>
#include <vector>

#include "c8mi/c8mi_tables.h"
#include "runtime.hpp"

namespace c8mi::em {

class DeviceTables {
 public:
  // Checks every descriptor against the array sizes it indexes (the
  // memory-safety part of validate_proposal_native_table,
  // src/kokkos/em/tables/ProposalNativeTable.cpp:64-160,234-420) and throws
  // std::invalid_argument on the first violation; then deep-copies.
  explicit DeviceTables(c8mi_native_table_set const& host);

  c8mi_native_table_set const& view() const noexcept { return view_; }
  std::uint64_t bytes() const noexcept { return storage_.bytes(); }

 private:
  DeviceBuffer<unsigned char> storage_;
  c8mi_native_table_set view_{};
};

// Host copy of a table set in the same packed layout (the host stepper's
// tables); needs no device.
class HostTables {
 public:
  explicit HostTables(c8mi_native_table_set const& host);
  HostTables(HostTables const&) = delete;
  HostTables& operator=(HostTables const&) = delete;

  c8mi_native_table_set const& view() const noexcept { return view_; }

 private:
  std::vector<unsigned char> storage_;
  c8mi_native_table_set view_{};
};

// Bounds checks alone, usable without a device.
void validate_table_bounds(c8mi_native_table_set const& host);

}  // namespace c8mi::em
