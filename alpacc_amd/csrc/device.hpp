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

// One DPP move of a 32-bit value: lanes whose source lane is outside the
// control's reach, or that the row / bank masks exclude, get `id`.
template <int CTRL, int ROW, int BANK>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t v, uint32_t id) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, CTRL, ROW, BANK, false);
}

// Inclusive wave64 scan by DPP, no LDS round trips: shifts by 1, 2, 3
// inside each row of 16 lanes, then by 4 and 8 (banks that have such a
// source), then the row broadcasts of lane 15 into rows 1 and 3 and of lane
// 31 into rows 2 and 3.  T is a tuple of 32-bit words moved one by one;
// op(a, b) joins an earlier lane's a with a later lane's b, id is its
// identity.
template <typename T, typename Mov, typename Op>
__device__ __forceinline__ T dpp_scan(T v, T id, Mov mov, Op op) {
  T x = op(mov.template operator()<0x111, 0xf, 0xf>(v, id), v);
  x = op(mov.template operator()<0x112, 0xf, 0xf>(v, id), x);
  x = op(mov.template operator()<0x113, 0xf, 0xf>(v, id), x);
  x = op(mov.template operator()<0x114, 0xf, 0xe>(x, id), x);
  x = op(mov.template operator()<0x118, 0xf, 0xc>(x, id), x);
  x = op(mov.template operator()<0x142, 0xa, 0xf>(x, id), x);
  x = op(mov.template operator()<0x143, 0xc, 0xf>(x, id), x);
  return x;
}

struct Mov1 {
  template <int C, int R, int B>
  __device__ __forceinline__ uint32_t operator()(uint32_t v, uint32_t id) const {
    return dpp_mov<C, R, B>(v, id);
  }
};
struct U32x2 {
  uint32_t a, b;
};
struct Mov2 {
  template <int C, int R, int B>
  __device__ __forceinline__ U32x2 operator()(U32x2 v, U32x2 id) const {
    return {dpp_mov<C, R, B>(v.a, id.a), dpp_mov<C, R, B>(v.b, id.b)};
  }
};

// Inclusive wave scan of a (count, max-position) pair.
__device__ __forceinline__ void wave_scan_add_max(uint32_t& cnt, uint32_t& mx) {
  const U32x2 r = dpp_scan(U32x2{cnt, mx}, U32x2{0u, 0u}, Mov2{}, [](U32x2 a, U32x2 b) {
    return U32x2{a.a + b.a, a.b > b.b ? a.b : b.b};
  });
  cnt = r.a;
  mx = r.b;
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
  return dpp_scan(v, 0u, Mov1{}, [](uint32_t a, uint32_t b) { return a + b; });
}

// Inclusive wave scan of the minimum.
__device__ __forceinline__ int32_t wave_scan_min32(int32_t v) {
  return (int32_t)dpp_scan((uint32_t)v, 0x7FFFFFFFu, Mov1{}, [](uint32_t a, uint32_t b) {
    return (uint32_t)((int32_t)a < (int32_t)b ? (int32_t)a : (int32_t)b);
  });
}

}  // namespace alpacc
