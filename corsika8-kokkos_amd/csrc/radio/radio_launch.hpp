// Host-callable launchers for the radio kernels (defined in radio_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "c8mi/c8mi_types.h"
#include "radio_device.hpp"

namespace c8mi::radio {

// Default threads per projection workgroup (one observer x one track chunk);
// 256, 512 and 1024 are instantiated.
inline constexpr std::uint32_t kProjectionBlock = 256;
// Dynamic LDS a projection workgroup may request (gfx950: 160 KiB per CU).
inline constexpr std::size_t kMaxDynamicLdsBytes = 144 * 1024;
// Default window tile: its dynamic LDS plus the kernel's static LDS stay within
// this budget, so three workgroups share a CU (160 KiB). A 4096-bin ZHS window
// in one 96 KiB tile (one workgroup per CU) measured 2x slower than two
// 2048-bin tiles; a 1024-bin CoREAS+ZHS window (48 KiB of bins) and a 2048-bin
// ZHS window keep one tile.
inline constexpr std::size_t kPreferredWorkgroupLdsBytes = 53 * 1024;
// Upper bound of the projection kernel's static LDS: the wave totals of the ZHS
// scan (16 x 8 B for 1024 threads), block counters and |increment| sums.
inline constexpr std::size_t kProjectionStaticLdsBytes = 256;
// Tracks per workgroup the default chunking gives a batch large enough for
// eight such chunks, at least (a smaller batch is cut into eight chunks).
inline constexpr std::uint64_t kMinimumDefaultChunk = 32768;
// XCDs of the chip: the projection grid maps chunk c to XCD c % kXcdCount.
inline constexpr std::uint64_t kXcdCount = 8;
// Bins of the default tile for `per_bin` int64 words per bin (3 per algorithm);
// a ZHS tile holds one more word per component row (its difference array).
inline std::uint64_t default_tile_capacity(std::size_t per_bin, bool zhs) {
  std::size_t const extra = zhs ? 3u : 0u;
  std::uint64_t const most = (kMaxDynamicLdsBytes / sizeof(long long) - extra) / per_bin;
  std::uint64_t const preferred =
      ((kPreferredWorkgroupLdsBytes - kProjectionStaticLdsBytes) / sizeof(long long) - extra) / per_bin;
  return preferred < most ? preferred : most;
}

// Launch-shape knobs of a projection (set_tuning): tracks per workgroup
// (0 = automatic), window tile limit in bins (0 = automatic), target
// workgroups per launch and threads per workgroup.
struct LaunchTuning {
  std::uint64_t chunk_tracks = 0;
  std::uint32_t tile_bins_limit = 0;
  std::uint32_t target_workgroups = 8192;
  std::uint32_t block_threads = 512;
};

// Fills the tile and chunk fields of `L` (tile_bins, bin_tiles, chunk_size,
// chunks) for `count` tracks on `observers` observers whose longest window has
// `max_bins` bins. Throws length_error when the grid exceeds 2^32 workgroups.
void shape_projection(ProjectionLaunch& L, bool coreas, bool zhs, std::uint32_t observers,
                      std::uint64_t max_bins, std::uint64_t count, LaunchTuning const& tuning);

// LDS bytes for one workgroup holding `tile_bins` bins of each enabled mode.
std::size_t projection_lds_bytes(bool coreas, bool zhs, std::uint32_t tile_bins);
// Workgroups launched for `launch` (chunks rounded up to whole XCD rounds).
std::uint32_t projection_grid(ProjectionLaunch const& launch);

// exact: the reference's FP64 operation order (bit-identical fixed-point
// sums); otherwise the fast FP64 form. Floating accumulation is always exact.
void launch_projection(ProjectionLaunch const& launch, bool coreas, bool zhs, bool fixed,
                       bool exact, std::uint32_t block, hipStream_t stream);

// Pooled showers (launch.segments set): `grid` workgroups, an upper bound on
// the items (*launch.item_count); int64 fixed point.
void launch_pooled_projection(ProjectionLaunch const& launch, bool coreas, bool zhs, bool exact,
                              std::uint32_t block, std::uint32_t grid, hipStream_t stream);

void launch_records_to_tracks(c8mi_lepton_transport_record const* records,
                              std::uint64_t count, double* tracks, std::uint64_t stride,
                              hipStream_t stream);

}  // namespace c8mi::radio
