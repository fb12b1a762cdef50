// DFA lexer for gfx950 in three launches.
//
// Semantics are those of the sequential reference lexer (backends/c/lexer.c:
// 39-97): greedy tokens with no backtracking, a boundary before byte i when
// the transition on byte i is a "producing" one, ignore tokens dropped, a kept
// token longer than length_t rejects the input, and the input is accepted iff
// the state after the last byte is accepting.
//
// The input is cut into 32 KiB tiles of 256 lanes x 128 bytes.
//   L1 lex_fold      every lane folds its 128 bytes into one endomorphism
//                    (LDS tables STEP[e][byte class]); a workgroup scan
//                    (compose table in L2) gives each lane its exclusive
//                    prefix inside the tile (u16, written out) and the tile
//                    its aggregate;
//   L2 lex_tile_scan one workgroup scans the tile aggregates;
//   L3 lex_emit      every lane starts from the DFA state
//                    ENDO_INIT[tile prefix . lane prefix], walks its bytes
//                    through the LDS token DFA (next | produce | token),
//                    marks boundaries and the terminals of closed tokens, and
//                    gets its output slot and the start of its first token
//                    from a block scan plus a decoupled lookback over
//                    (kept-token count, last boundary) -- sum and max, so a
//                    whole wave inspects 64 predecessor tiles per probe --
//                    then writes the tokens as structure-of-arrays.
#include "device.hpp"
#include "lexer.hpp"

namespace alpacc {

namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int WG = LEX_WG;
constexpr int IPT = LEX_IPT;
constexpr int NW = WG / WAVE;
constexpr uint32_t TILE = (uint32_t)WG * IPT;

// Load the lane's IPT bytes (zero past the end) as IPT/4 little-endian words.
__device__ __forceinline__ void load_lane(const LexParams& p, uint64_t base, uint32_t* w) {
  const uint64_t n = p.n;
  if (base + IPT <= n && p.aligned16) {
    const u32x4* src = reinterpret_cast<const u32x4*>(p.in + base);
#pragma unroll
    for (int v = 0; v < IPT / 16; ++v) {
      u32x4 q = __builtin_nontemporal_load(src + v);
      w[4 * v + 0] = q.x;
      w[4 * v + 1] = q.y;
      w[4 * v + 2] = q.z;
      w[4 * v + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int j = 0; j < IPT / 4; ++j) {
      uint32_t x = 0;
      for (int b = 0; b < 4; ++b) {
        uint64_t pos = base + 4 * j + b;
        if (pos < n) x |= (uint32_t)p.in[pos] << (8 * b);
      }
      w[j] = x;
    }
  }
}

__device__ __forceinline__ uint32_t byte_at(const uint32_t* w, int j) { return (w[j >> 2] >> (8 * (j & 3))) & 0xFFu; }

// ------------------------------------------------------------------- L1

__global__ __launch_bounds__(WG) void lex_fold(LexParams p) {
  extern __shared__ __align__(16) uint8_t lds[];
  uint8_t* s_class = lds;
  uint16_t* s_step = reinterpret_cast<uint16_t*>(lds + 256);
  __shared__ uint32_t s_wave[NW];
  const uint32_t tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
  for (int i = tid; i < 256; i += WG) s_class[i] = p.byte_class[i];
  for (int i = tid; i < p.nendo * p.nclass; i += WG) s_step[i] = p.step[i];
  __syncthreads();
  const uint16_t* __restrict__ compose = p.compose;
  const uint32_t nendo = p.nendo, nclass = p.nclass;
  auto comp = [&](uint32_t a, uint32_t b) -> uint32_t { return compose[a * nendo + b]; };

  for (uint32_t tile = blockIdx.x; tile < p.ntiles; tile += gridDim.x) {
    const uint64_t base = (uint64_t)tile * TILE + (uint64_t)tid * IPT;
    uint32_t w[IPT / 4];
    load_lane(p, base, w);
    const uint32_t valid = base >= p.n ? 0u : (base + IPT <= p.n ? (uint32_t)IPT : (uint32_t)(p.n - base));
    uint32_t e = 0;  // identity
    // fully unrolled: the byte registers are indexed with constants
    if (valid == IPT) {
#pragma unroll
      for (int j = 0; j < IPT; ++j) e = s_step[e * nclass + s_class[byte_at(w, j)]];
    } else {
#pragma unroll
      for (int j = 0; j < IPT; ++j)
        if ((uint32_t)j < valid) e = s_step[e * nclass + s_class[byte_at(w, j)]];
    }
    uint32_t inc = e;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
      uint32_t o = __shfl_up(inc, d, WAVE);
      if (lane >= (uint32_t)d) inc = comp(o, inc);
    }
    if (lane == WAVE - 1) s_wave[wid] = inc;
    __syncthreads();
    uint32_t wpre = 0;
    for (uint32_t i = 0; i < wid; ++i) wpre = comp(wpre, s_wave[i]);
    uint32_t prev = __shfl_up(inc, 1, WAVE);
    uint32_t excl = lane == 0 ? wpre : comp(wpre, prev);
    p.lane_prefix[(uint64_t)tile * WG + tid] = (uint16_t)excl;
    if (tid == WG - 1) p.tile_endo[tile] = comp(wpre, inc);
    __syncthreads();
  }
}

// ------------------------------------------------------------------- L2

// Exclusive scan of the tile endomorphisms by one workgroup of 1024 lanes:
// each lane folds a contiguous run, a wave scan and a sequential pass over
// the 16 wave totals join the runs, then each lane rewrites its run.
__global__ __launch_bounds__(1024) void lex_tile_scan(LexParams p) {
  __shared__ uint32_t s_wave[16];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const uint16_t* __restrict__ compose = p.compose;
  const uint32_t nendo = p.nendo;
  auto comp = [&](uint32_t a, uint32_t b) -> uint32_t { return compose[a * nendo + b]; };
  const uint32_t nt = p.ntiles;
  const uint32_t per = (nt + 1023) / 1024;
  const uint32_t lo = tid * per < nt ? tid * per : nt, hi = lo + per < nt ? lo + per : nt;
  uint32_t run = 0;
  for (uint32_t t = lo; t < hi; ++t) run = comp(run, p.tile_endo[t]);
  uint32_t inc = run;
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    uint32_t o = __shfl_up(inc, d, WAVE);
    if (lane >= (uint32_t)d) inc = comp(o, inc);
  }
  if (lane == 63) s_wave[wid] = inc;
  __syncthreads();
  uint32_t wpre = 0;
  for (uint32_t i = 0; i < wid; ++i) wpre = comp(wpre, s_wave[i]);
  uint32_t prev = __shfl_up(inc, 1, WAVE);
  uint32_t acc = lane == 0 ? wpre : comp(wpre, prev);
  for (uint32_t t = lo; t < hi; ++t) {
    uint32_t e = p.tile_endo[t];
    p.tile_prefix[t] = (uint16_t)acc;
    acc = comp(acc, e);
  }
}

// ------------------------------------------------------------------- L3

__global__ __launch_bounds__(WG) void lex_emit(LexParams p) {
  extern __shared__ __align__(16) uint8_t lds[];
  uint16_t* s_delta = reinterpret_cast<uint16_t*>(lds);  // nstate*256
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_wave_cnt[NW], s_wave_last[NW];
  __shared__ uint32_t s_pre_cnt, s_pre_last;
  const uint32_t tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
  for (int i = tid; i < p.nstate * 256; i += WG) s_delta[i] = p.delta[i];
  const uint16_t* __restrict__ compose = p.compose;
  const uint32_t nendo = p.nendo;

  for (;;) {
    if (tid == 0) s_tile = atomicAdd(p.ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= p.ntiles) break;
    const uint64_t n = p.n;
    const uint64_t base = (uint64_t)tile * TILE + (uint64_t)tid * IPT;
    const uint32_t valid = base >= n ? 0u : (base + IPT <= n ? (uint32_t)IPT : (uint32_t)(n - base));
    uint32_t w[IPT / 4];
    load_lane(p, base, w);
    uint32_t pre = compose[(uint32_t)p.tile_prefix[tile] * nendo + p.lane_prefix[(uint64_t)tile * WG + tid]];
    uint32_t s = p.endo_init[pre];

    // walk: boundary / kept masks and the terminals of the closed tokens
    uint32_t bmask[IPT / 32], kmask[IPT / 32], terms[IPT / 4];
#pragma unroll
    for (int j = 0; j < IPT / 32; ++j) bmask[j] = kmask[j] = 0;
#pragma unroll
    for (int j = 0; j < IPT / 4; ++j) terms[j] = 0;
    uint32_t cnt = 0, last = 0;
    const int ignore = p.ignore;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      if ((uint32_t)j < valid) {
        uint32_t x = s_delta[s * 256 + byte_at(w, j)];
        if (x & 0x100u) {
          uint32_t t = x >> 9;
          bmask[j >> 5] |= 1u << (j & 31);
          last = (uint32_t)(base + j);
          if ((int)t != ignore) {
            kmask[j >> 5] |= 1u << (j & 31);
            terms[j >> 2] |= t << (8 * (j & 3));
            ++cnt;
          }
        }
        s = x & 0xFFu;
      }
    }
    const bool owns_end = valid > 0 && base + valid == n;
    uint32_t fin_t = 0;
    bool fin_kept = false;
    if (owns_end) {
      fin_t = s_delta[s * 256] >> 9;
      if ((int)fin_t == p.dead) {
        atomicOr(p.flags, LEX_REJECT_FINAL);
      } else if ((int)fin_t != ignore) {
        fin_kept = true;
        ++cnt;
      }
    }

    // block scan of (count, last boundary)
    uint32_t icnt = cnt, ilast = last;
    wave_scan_add_max(icnt, ilast);
    if (lane == WAVE - 1) {
      s_wave_cnt[wid] = icnt;
      s_wave_last[wid] = ilast;
    }
    __syncthreads();
    uint32_t wc = 0, wl = 0, tc = 0, tl = 0;
    for (uint32_t i = 0; i < NW; ++i) {
      if (i < wid) {
        wc += s_wave_cnt[i];
        wl = wl > s_wave_last[i] ? wl : s_wave_last[i];
      }
      tc += s_wave_cnt[i];
      tl = tl > s_wave_last[i] ? tl : s_wave_last[i];
    }
    uint32_t pc = __shfl_up(icnt, 1, WAVE), pl = __shfl_up(ilast, 1, WAVE);
    uint32_t ecnt = wc + (lane ? pc : 0u);
    uint32_t elast = lane ? (wl > pl ? wl : pl) : wl;

    // decoupled lookback by wave 0: 64 predecessor tiles per probe
    if (wid == 0) {
      const uint64_t agg = (uint64_t)tc | ((uint64_t)tl << 31);
      uint32_t xc = 0, xl = 0;
      if (tile == 0) {
        if (lane == 0) publish64(p.count_status, ST64_PRE | agg);
      } else {
        if (lane == 0) publish64(p.count_status + tile, ST64_AGG | agg);
        int64_t top = (int64_t)tile - 1;
        for (;;) {
          int64_t idx = top - (int64_t)lane;
          uint64_t st = idx >= 0 ? observe64(p.count_status + idx) : ST64_PRE;
          uint64_t fl = st & ST64_MASK;
          uint64_t zero = __ballot(fl == 0);
          uint64_t prem = __ballot(fl == ST64_PRE);
          int first_pre = prem ? __builtin_ctzll(prem) : WAVE;
          int first_zero = zero ? __builtin_ctzll(zero) : WAVE;
          if (first_zero < first_pre) {  // a nearer tile has not published yet
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          uint32_t c = (int)lane <= first_pre && idx >= 0 ? (uint32_t)(st & 0x7FFFFFFFu) : 0u;
          uint32_t l = (int)lane <= first_pre && idx >= 0 ? (uint32_t)((st >> 31) & 0x7FFFFFFFu) : 0u;
#pragma unroll
          for (int d = 32; d >= 1; d >>= 1) {
            c += __shfl_xor(c, d, WAVE);
            uint32_t o = __shfl_xor(l, d, WAVE);
            l = l > o ? l : o;
          }
          xc += c;
          xl = xl > l ? xl : l;
          if (first_pre < WAVE) break;
          top -= WAVE;
        }
        if (lane == 0) {
          uint32_t ic = xc + tc, il = xl > tl ? xl : tl;
          publish64(p.count_status + tile, ST64_PRE | (uint64_t)ic | ((uint64_t)il << 31));
        }
      }
      if (lane == 0) {
        s_pre_cnt = xc;
        s_pre_last = xl;
      }
    }
    __syncthreads();

    // emit
    uint64_t slot = (uint64_t)s_pre_cnt + ecnt;
    uint64_t start = s_pre_last > elast ? s_pre_last : elast;
    const uint64_t max_len = p.max_len;
    bool overflow = false;
    // unrolled so the packed terminal registers are indexed with constants
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      if ((bmask[j >> 5] >> (j & 31)) & 1u) {
        uint64_t pos = base + j;
        if ((kmask[j >> 5] >> (j & 31)) & 1u) {
          uint32_t t = (terms[j >> 2] >> (8 * (j & 3))) & 0xFFu;
          uint64_t len = pos - start;
          overflow |= len > max_len;
          p.terminals[slot] = (uint8_t)t;
          p.starts[slot] = (int64_t)start;
          store_length(p, slot, len);
          ++slot;
        }
        start = pos;
      }
    }
    if (fin_kept) {
      uint64_t len = n - start;
      overflow |= len > max_len;
      p.terminals[slot] = (uint8_t)fin_t;
      p.starts[slot] = (int64_t)start;
      store_length(p, slot, len);
      ++slot;
    }
    if (overflow) atomicOr(p.flags, LEX_REJECT_OVERFLOW);
    if (owns_end) p.count[0] = slot;
    __syncthreads();
  }
}

template <typename K>
hipError_t occupancy_grid(K k, size_t lds, int* grid) {
  int dev = 0, cus = 0, per = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(k), WG, lds);
  if (e != hipSuccess) return e;
  *grid = cus * (per < 1 ? 1 : per);
  return hipSuccess;
}

}  // namespace

size_t lex_fold_lds(const LexParams& p) { return 256 + (size_t)p.nendo * p.nclass * 2; }
size_t lex_emit_lds(const LexParams& p) { return (size_t)p.nstate * 256 * 2; }
uint32_t lex_tile_bytes() { return TILE; }

hipError_t lex_grid_size(const LexParams& p, int* fold_grid, int* emit_grid) {
  hipError_t e = occupancy_grid(lex_fold, lex_fold_lds(p), fold_grid);
  if (e != hipSuccess) return e;
  return occupancy_grid(lex_emit, lex_emit_lds(p), emit_grid);
}

hipError_t lex_launch(const LexParams& p, int fold_grid, int emit_grid, hipStream_t st) {
  const uint32_t nt = p.ntiles;
  int g1 = fold_grid < (int)nt ? fold_grid : (int)nt;
  int g3 = emit_grid < (int)nt ? emit_grid : (int)nt;
  hipLaunchKernelGGL(lex_fold, dim3(g1), dim3(WG), lex_fold_lds(p), st, p);
  hipLaunchKernelGGL(lex_tile_scan, dim3(1), dim3(1024), 0, st, p);
  hipLaunchKernelGGL(lex_emit, dim3(g3), dim3(WG), lex_emit_lds(p), st, p);
  return hipGetLastError();
}

}  // namespace alpacc
