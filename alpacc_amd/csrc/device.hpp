// Device helpers shared by the lexer and parser kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace alpacc {

constexpr int WAVE = 64;

// Status words of the decoupled lookback: the flag lives in the top two bits
// of the same word as the payload, so one atomic publishes both and no other
// memory has to be ordered around it.  The accesses are therefore relaxed
// (monotonic) at agent scope: coherent across the XCDs' L2s (sc1), but with
// no release/acquire fences -- on gfx950 those are a writeback and an
// invalidate of the whole L2, paid on every publish and every probe.
constexpr uint32_t ST32_AGG = 1u << 30;
constexpr uint32_t ST32_PRE = 2u << 30;
constexpr uint32_t ST32_MASK = 3u << 30;
constexpr uint64_t ST64_AGG = 1ull << 62;
constexpr uint64_t ST64_PRE = 2ull << 62;
constexpr uint64_t ST64_MASK = 3ull << 62;

__device__ __forceinline__ void publish32(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t observe32(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void publish64(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t observe64(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Inclusive wave scan of a (count, max-position) pair.
__device__ __forceinline__ void wave_scan_add_max(uint32_t& cnt, uint32_t& mx) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    uint32_t c = __shfl_up(cnt, d, WAVE);
    uint32_t m = __shfl_up(mx, d, WAVE);
    if (lane >= (uint32_t)d) {
      cnt += c;
      mx = mx > m ? mx : m;
    }
  }
}

__device__ __forceinline__ uint64_t wave_scan_add64(uint64_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    uint64_t o = __shfl_up(v, d, WAVE);
    if (lane >= (uint32_t)d) v += o;
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_scan_add32(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    uint32_t o = __shfl_up(v, d, WAVE);
    if (lane >= (uint32_t)d) v += o;
  }
  return v;
}

}  // namespace alpacc
