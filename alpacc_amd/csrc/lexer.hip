// Single-pass DFA lexer for gfx950.
//
// Semantics are those of the sequential reference lexer (backends/c/lexer.c:
// 39-97): greedy tokens with no backtracking, a boundary before byte i when
// the transition on byte i is a "producing" one, ignore tokens dropped, a kept
// token longer than length_t rejects the input, and the input is accepted iff
// the state after the last byte is accepting.
//
// Work decomposition (one persistent kernel, tiles handed out by an atomic
// ticket so a tile only ever waits on tiles that were handed out earlier):
//   A  every lane folds its IPT bytes into one endomorphism id through the
//      LDS table STEP[e][byte class];
//   B  workgroup scan of the endomorphisms (compose table in L2) plus a
//      decoupled lookback over tiles gives each lane the DFA state at its
//      first byte (ENDO_INIT[prefix]);
//   C  every lane walks its bytes through the LDS token DFA, marking token
//      boundaries and the terminals of the tokens they close; a second scan
//      and lookback over (kept-token count, last boundary) gives each lane its
//      output slot and the start of its first token;
//   D  tokens are written as structure-of-arrays (terminal, start, length).
#include "device.hpp"
#include "lexer.hpp"

namespace alpacc {

namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int WG, int IPT>
__global__ __launch_bounds__(WG) void lex_tiles(LexParams p) {
  static_assert(IPT % 16 == 0 && IPT <= 32, "IPT must be 16 or 32");
  constexpr int NW = WG / WAVE;
  constexpr uint32_t TILE = WG * IPT;
  extern __shared__ __align__(16) uint8_t lds[];
  uint8_t* s_class = lds;                                          // 256
  uint16_t* s_delta = reinterpret_cast<uint16_t*>(lds + 256);      // nstate*256
  uint16_t* s_step = s_delta + p.nstate * 256;                     // nendo*nclass
  uint8_t* s_init = reinterpret_cast<uint8_t*>(s_step + p.nendo * p.nclass);  // nendo

  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_wave_endo[NW];
  __shared__ uint32_t s_wave_cnt[NW];
  __shared__ uint32_t s_wave_last[NW];
  __shared__ uint32_t s_agg_endo;
  __shared__ uint32_t s_pre_endo;
  __shared__ uint32_t s_pre_cnt;
  __shared__ uint32_t s_pre_last;

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & (WAVE - 1);
  const uint32_t wid = tid / WAVE;

  // Stage the lexer tables once per persistent workgroup.
  for (int i = tid; i < 256; i += WG) s_class[i] = p.byte_class[i];
  for (int i = tid; i < p.nstate * 256; i += WG) s_delta[i] = p.delta[i];
  for (int i = tid; i < p.nendo * p.nclass; i += WG) s_step[i] = p.step[i];
  for (int i = tid; i < p.nendo; i += WG) s_init[i] = p.endo_init[i];

  const uint16_t* __restrict__ compose = p.compose;
  const uint32_t nendo = p.nendo;
  auto comp = [&](uint32_t a, uint32_t b) -> uint32_t { return compose[a * nendo + b]; };

  for (;;) {
    if (tid == 0) s_tile = atomicAdd(p.ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= p.ntiles) break;

    const uint64_t n = p.n;
    const uint64_t base = (uint64_t)tile * TILE + (uint64_t)tid * IPT;
    const bool full = base + IPT <= n;

    // ---- load IPT bytes --------------------------------------------------
    uint32_t w[IPT / 4];
    if (full && p.aligned16) {
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
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          uint64_t pos = base + 4 * j + b;
          if (pos < n) x |= (uint32_t)p.in[pos] << (8 * b);
        }
        w[j] = x;
      }
    }
    const uint32_t valid = full ? IPT : (base < n ? (uint32_t)(n - base) : 0u);

    // ---- A: fold the lane's bytes into one endomorphism -----------------
    uint32_t e = 0;  // identity
    if (full) {
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        uint32_t b = (w[j >> 2] >> (8 * (j & 3))) & 0xFF;
        e = s_step[e * p.nclass + s_class[b]];
      }
    } else {
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        uint32_t b = (w[j >> 2] >> (8 * (j & 3))) & 0xFF;
        if ((uint32_t)j < valid) e = s_step[e * p.nclass + s_class[b]];
      }
    }

    // ---- B: scan of endomorphisms + lookback ----------------------------
    uint32_t inc = e;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
      uint32_t o = __shfl_up(inc, d, WAVE);
      if (lane >= (uint32_t)d) inc = comp(o, inc);
    }
    if (lane == WAVE - 1) s_wave_endo[wid] = inc;
    __syncthreads();
    uint32_t wpre = 0;
    for (uint32_t i = 0; i < wid; ++i) wpre = comp(wpre, s_wave_endo[i]);
    uint32_t prev = __shfl_up(inc, 1, WAVE);
    uint32_t excl = lane == 0 ? wpre : comp(wpre, prev);
    if (tid == WG - 1) s_agg_endo = comp(wpre, inc);
    __syncthreads();
    if (tid == 0) {
      const uint32_t agg = s_agg_endo;
      uint32_t pre = 0;
      if (tile == 0) {
        publish32(p.endo_status + tile, ST32_PRE | agg);
      } else {
        publish32(p.endo_status + tile, ST32_AGG | agg);
        uint32_t acc = 0;
        int64_t j = (int64_t)tile - 1;
        for (;;) {
          uint32_t st = observe32(p.endo_status + j);
          uint32_t fl = st & ST32_MASK;
          if (fl == 0) {
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          acc = comp(st & 0xFFFFu, acc);
          if (fl == ST32_PRE) break;
          --j;
        }
        pre = acc;
        publish32(p.endo_status + tile, ST32_PRE | comp(pre, agg));
      }
      s_pre_endo = pre;
    }
    __syncthreads();
    uint32_t s = s_init[comp(s_pre_endo, excl)];

    // ---- C: walk the token DFA ------------------------------------------
    uint32_t bmask = 0, kmask = 0, cnt = 0, last = 0;
    uint32_t terms[IPT / 4];
#pragma unroll
    for (int j = 0; j < IPT / 4; ++j) terms[j] = 0;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      if (full || (uint32_t)j < valid) {
        uint32_t b = (w[j >> 2] >> (8 * (j & 3))) & 0xFF;
        uint32_t x = s_delta[s * 256 + b];
        if (x & 0x100u) {
          uint32_t t = x >> 9;
          bmask |= 1u << j;
          last = (uint32_t)(base + j);
          if ((int)t != p.ignore) {
            kmask |= 1u << j;
            terms[j >> 2] |= t << (8 * (j & 3));
            ++cnt;
          }
        }
        s = x & 0xFFu;
      }
    }
    // The lane holding the last input byte closes the final token.
    const bool owns_end = valid > 0 && base + valid == n;
    uint32_t fin_t = 0;
    bool fin_kept = false;
    if (owns_end) {
      fin_t = s_delta[s * 256] >> 9;
      if ((int)fin_t == p.dead) {
        atomicOr(p.flags, LEX_REJECT_FINAL);
      } else if ((int)fin_t != p.ignore) {
        fin_kept = true;
        ++cnt;
      }
    }

    // scan of (count, last boundary) + lookback
    uint32_t icnt = cnt, ilast = last;
    wave_scan_add_max(icnt, ilast);
    if (lane == WAVE - 1) {
      s_wave_cnt[wid] = icnt;
      s_wave_last[wid] = ilast;
    }
    __syncthreads();
    uint32_t wc = 0, wl = 0;
    for (uint32_t i = 0; i < wid; ++i) {
      wc += s_wave_cnt[i];
      wl = wl > s_wave_last[i] ? wl : s_wave_last[i];
    }
    uint32_t pc = __shfl_up(icnt, 1, WAVE), pl = __shfl_up(ilast, 1, WAVE);
    uint32_t ecnt = wc + (lane ? pc : 0u);
    uint32_t elast = lane ? (wl > pl ? wl : pl) : wl;
    if (tid == WG - 1) {
      uint32_t acnt = wc + icnt, alast = wl > ilast ? wl : ilast;
      uint64_t agg = (uint64_t)acnt | ((uint64_t)alast << 31);
      uint32_t precnt = 0, prelast = 0;
      if (tile == 0) {
        publish64(p.count_status + tile, ST64_PRE | agg);
      } else {
        publish64(p.count_status + tile, ST64_AGG | agg);
        int64_t j = (int64_t)tile - 1;
        for (;;) {
          uint64_t st = observe64(p.count_status + j);
          uint64_t fl = st & ST64_MASK;
          if (fl == 0) {
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          uint32_t c = (uint32_t)(st & 0x7FFFFFFFu);
          uint32_t l = (uint32_t)((st >> 31) & 0x7FFFFFFFu);
          precnt += c;
          prelast = prelast > l ? prelast : l;
          if (fl == ST64_PRE) break;
          --j;
        }
        uint32_t tc = precnt + acnt, tl = prelast > alast ? prelast : alast;
        publish64(p.count_status + tile, ST64_PRE | (uint64_t)tc | ((uint64_t)tl << 31));
      }
      s_pre_cnt = precnt;
      s_pre_last = prelast;
    }
    __syncthreads();

    // ---- D: emit ---------------------------------------------------------
    uint64_t slot = (uint64_t)s_pre_cnt + ecnt;
    uint64_t start = s_pre_last > elast ? s_pre_last : elast;
    const uint64_t max_len = p.max_len;
    bool overflow = false;
    // unrolled so the packed terminal registers are indexed with constants
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      if ((bmask >> j) & 1u) {
        uint64_t pos = base + j;
        if ((kmask >> j) & 1u) {
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

}  // namespace

template <int WG, int IPT>
static hipError_t launch_lex(const LexParams& p, int grid, size_t lds, hipStream_t st) {
  void (*k)(LexParams) = lex_tiles<WG, IPT>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(grid), dim3(WG), lds, st, p);
  return hipGetLastError();
}

size_t lex_lds_bytes(const LexParams& p) {
  return 256 + (size_t)p.nstate * 256 * 2 + (size_t)p.nendo * p.nclass * 2 + p.nendo;
}

uint32_t lex_tile_bytes() { return LEX_WG * LEX_IPT; }

hipError_t lex_grid_size(const LexParams& p, int* grid) {
  int dev = 0, cus = 0, per = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  void (*k)(LexParams) = lex_tiles<LEX_WG, LEX_IPT>;
  size_t lds = lex_lds_bytes(p);
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(k), LEX_WG, lds);
  if (e != hipSuccess) return e;
  if (per < 1) per = 1;
  *grid = cus * per;
  return hipSuccess;
}

hipError_t lex_launch(const LexParams& p, int grid, hipStream_t st) {
  return launch_lex<LEX_WG, LEX_IPT>(p, grid, lex_lds_bytes(p), st);
}

}  // namespace alpacc
