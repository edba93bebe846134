// Format-v7 table sets on disk: the reference's canonical encoding as payload,
// its SHA-256 as the content hash (calculate_proposal_native_hash,
// src/kokkos/em/tables/ProposalNativeTable.cpp:570-666). Host only.
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <vector>

#include "c8mi/c8mi_tables.h"

namespace c8mi::em {

using Digest = std::array<std::uint8_t, 32>;

// SHA-256 (FIPS 180-4).
Digest sha256(std::uint8_t const* data, std::size_t size);

// The content hash of a table set with its metadata.
Digest native_table_hash(c8mi_native_table_set const& set, c8mi_native_table_meta const& meta);

// A decoded table file: owning arrays and the set / metadata views into them.
class TableFile {
 public:
  explicit TableFile(std::vector<std::uint8_t> const& file);  // std::invalid_argument on a bad file
  c8mi_native_table_set const& set() const noexcept { return set_; }
  c8mi_native_table_meta const& meta() const noexcept { return meta_; }
  Digest const& digest() const noexcept { return digest_; }

 private:
  std::vector<c8mi_native_interaction_identity> identities_;
  std::vector<c8mi_native_dndx_column> dndx_;
  std::vector<c8mi_native_total_rate_column> totals_;
  std::vector<std::uint32_t> selection_;
  std::vector<c8mi_native_dedx_column> dedx_;
  std::vector<std::uint32_t> accumulation_;
  std::vector<c8mi_native_utility_column> utility_;
  std::vector<double> bicubic_[4];
  std::vector<double> polynomial_;
  std::vector<double> cubic_[2];
  c8mi_native_table_set set_{};
  c8mi_native_table_meta meta_{};
  Digest digest_{};
};

std::unique_ptr<TableFile> read_native_table(char const* path);
// Writes through a temporary file next to `path`, read back and verified
// before it is renamed over `path`.
void write_native_table(char const* path, c8mi_native_table_set const& set, c8mi_native_table_meta const& meta);

}  // namespace c8mi::em
