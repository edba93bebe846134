// Counter-based Philox4x32-10 stream of the EM path
// (corsika/detail/kokkos/em/random/Philox.inl:15-57). The counter is
// (historyId, stepId); the key is splitmix64 chained over seed, showerId,
// processId and drawId. A draw depends only on these six identities, never on
// queue position, so results do not depend on how particles are batched.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace c8mi::em {

struct RandomKey {
  std::uint64_t seed;
  std::uint64_t shower;
  std::uint64_t history;
  std::uint64_t step;
  std::uint32_t process;
  std::uint64_t draw;
};

__host__ __device__ inline std::uint64_t splitmix64(std::uint64_t v) {
  v += 0x9e3779b97f4a7c15ULL;
  v = (v ^ (v >> 30U)) * 0xbf58476d1ce4e5b9ULL;
  v = (v ^ (v >> 27U)) * 0x94d049bb133111ebULL;
  return v ^ (v >> 31U);
}

// Ten Philox rounds (multipliers 0xD2511F53 / 0xCD9E8D57, key increments
// 0x9E3779B9 / 0xBB67AE85): v_mul_hi_u32 / v_mul_lo_u32 pairs on the VALU.
__host__ __device__ inline void philox4x32_10(std::uint32_t c[4], std::uint32_t k0,
                                              std::uint32_t k1) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    std::uint64_t const p0 = static_cast<std::uint64_t>(0xD2511F53u) * c[0];
    std::uint64_t const p1 = static_cast<std::uint64_t>(0xCD9E8D57u) * c[2];
    std::uint32_t const n0 = static_cast<std::uint32_t>(p1 >> 32) ^ c[1] ^ k0;
    std::uint32_t const n2 = static_cast<std::uint32_t>(p0 >> 32) ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = static_cast<std::uint32_t>(p1);
    c[2] = n2;
    c[3] = static_cast<std::uint32_t>(p0);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

__host__ __device__ inline void random_words(RandomKey const& key, std::uint32_t out[4]) {
  out[0] = static_cast<std::uint32_t>(key.history);
  out[1] = static_cast<std::uint32_t>(key.history >> 32U);
  out[2] = static_cast<std::uint32_t>(key.step);
  out[3] = static_cast<std::uint32_t>(key.step >> 32U);
  std::uint64_t m = splitmix64(key.seed);
  m = splitmix64(m ^ key.shower);
  m = splitmix64(m ^ static_cast<std::uint64_t>(key.process));
  m = splitmix64(m ^ key.draw);
  philox4x32_10(out, static_cast<std::uint32_t>(m), static_cast<std::uint32_t>(m >> 32U));
}

// (word + 0.5) / 2^32: never 0 or 1.
__host__ __device__ inline double uniform_open01(RandomKey const& key, std::uint32_t lane = 0) {
  std::uint32_t w[4];
  random_words(key, w);
  return (static_cast<double>(w[lane & 3U]) + 0.5) * (1. / 4294967296.);
}

// Random-process domains (corsika/kokkos/em/random/RandomDomains.hpp:16-31).
inline constexpr std::uint32_t kDistanceProcess = 0x454d0001U;
inline constexpr std::uint32_t kColumnProcess = 0x454d0002U;
inline constexpr std::uint32_t kScatteringProcess = 0x454d0003U;
inline constexpr std::uint32_t kSelectionProcess = 0x454d0004U;
inline constexpr std::uint32_t kMuonDecayProcess = 0x4d554445U;

}  // namespace c8mi::em
