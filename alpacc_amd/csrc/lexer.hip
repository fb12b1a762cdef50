// DFA lexer for gfx950 in three launches.
//
// Semantics are those of the sequential reference lexer (backends/c/lexer.c:
// 39-97): greedy tokens with no backtracking, a boundary before byte i when
// the transition on byte i is a "producing" one, ignore tokens dropped, a kept
// token longer than length_t rejects the input, and the input is accepted iff
// the state after the last byte is accepting.
//
// A launch lexes a batch of inputs packed back to back (a single input is a
// batch of one).  Every input owns whole tiles of 256 lanes x 128 bytes; a
// tile's first byte need not be aligned (loads are dword-aligned and bytes
// are realigned in registers).
//   L1 lex_fold      every lane folds its 128 bytes into one endomorphism
//                    (LDS tables STEP[e][byte class]); a workgroup scan
//                    (compose table in L2) gives each lane its exclusive
//                    prefix inside the tile (u16, written out) and the tile
//                    its aggregate;
//   L2 lex_tile_scan one workgroup scans the tile aggregates, restarting at
//                    the first tile of every input;
//   L3 lex_emit      every lane starts from the DFA state
//                    ENDO_INIT[tile prefix . lane prefix] and walks its bytes
//                    through the LDS token DFA (next | produce | kept),
//                    counting kept tokens and its last boundary; a block scan
//                    plus a decoupled lookback over (kept-token count, last
//                    boundary) -- sum and max, so a whole wave inspects 64
//                    predecessor tiles per probe -- gives the output slot and
//                    the start of every lane's first token (never before its
//                    input's first byte); a second walk stages (terminal,
//                    start, end) in LDS by slot, and the workgroup writes the
//                    structure-of-arrays output with consecutive lanes on
//                    consecutive slots.
#include "device.hpp"
#include "lexer.hpp"

namespace alpacc {

namespace {

// 16 bytes loaded from a dword-aligned address.
typedef uint32_t u32x4a4 __attribute__((ext_vector_type(4), aligned(4)));

constexpr int WG = LEX_WG;
constexpr int IPT = LEX_IPT;
constexpr int NW = WG / WAVE;
constexpr uint32_t TILE = LEX_TILE;

// A tile, its input and whether its bytes all belong to that input.
struct TileSpan {
  uint64_t tbase;   // offset of the tile's first byte in the packed inputs
  uint64_t seg_lo;  // its input: [seg_lo, seg_hi)
  uint64_t seg_hi;
  uint32_t seg;
  uint32_t first;   // the input's first tile
  uint32_t full;    // TILE bytes of the input
  uint32_t shift;   // byte misalignment of the tile start (the same for every lane)
};

__device__ __forceinline__ TileSpan tile_span(const LexParams& p, uint32_t tile) {
  TileSpan t;
  const uint32_t seg = p.tile_seg[tile];
  const uint32_t tf = p.tile_first[seg];
  t.seg = seg;
  t.first = tf == tile;
  t.seg_lo = p.seg_off[seg];
  t.seg_hi = p.seg_off[seg + 1];
  t.tbase = t.seg_lo + (uint64_t)(tile - tf) * TILE;
  t.full = t.tbase + TILE <= t.seg_hi;
  t.shift = (uint32_t)((uintptr_t)(p.in + t.tbase) & 3u);
  return t;
}

// The dwords holding the lane's IPT bytes of a full tile: a[0] holds its
// first byte at byte `shift`, a[IPT / 4] is read only when the tile is
// misaligned (it holds the lane's last byte), so no byte outside the dwords
// of the lane's own bytes is touched.
__device__ __forceinline__ void load_lane(const LexParams& p, uint64_t base, uint32_t shift, uint32_t* a) {
  const u32x4a4* src = reinterpret_cast<const u32x4a4*>(p.in + base - shift);
#pragma unroll
  for (int v = 0; v < IPT / 16; ++v) {
    const u32x4a4 q = src[v];
    a[4 * v + 0] = q.x;
    a[4 * v + 1] = q.y;
    a[4 * v + 2] = q.z;
    a[4 * v + 3] = q.w;
  }
  a[IPT / 4] = shift ? reinterpret_cast<const uint32_t*>(src)[IPT / 4] : 0u;
}

// Bytes 4k..4k+3 of the lane.  The empty asm makes the realignment wait
// for `dep` (the walk's state), so the 32 realigned words are formed as the
// walk reaches them instead of all at once next to the raw ones.
__device__ __forceinline__ uint32_t lane_word(const uint32_t* a, int k, uint32_t shift, uint32_t dep) {
  uint32_t lo = a[k], hi = a[k + 1];
  asm volatile("" : "+v"(lo), "+v"(hi) : "v"(dep));
  return __builtin_amdgcn_alignbyte(hi, lo, shift);
}

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
    const TileSpan ts = tile_span(p, tile);
    const uint64_t base = ts.tbase + (uint64_t)tid * IPT;
    uint32_t e = 0;  // identity
    // full tiles: fully unrolled (the byte registers are indexed with
    // constants); an input's last tile reads its bytes one by one
    if (ts.full) {
      uint32_t a[IPT / 4 + 1];
      load_lane(p, base, ts.shift, a);
#pragma unroll
      for (int k = 0; k < IPT / 4; ++k) {
        const uint32_t wk = lane_word(a, k, ts.shift, e);
#pragma unroll
        for (int b = 0; b < 4; ++b) e = s_step[e * nclass + s_class[(wk >> (8 * b)) & 0xFFu]];
      }
    } else {
      const uint32_t valid = base >= ts.seg_hi ? 0u : (uint32_t)min((uint64_t)IPT, ts.seg_hi - base);
      for (uint32_t j = 0; j < valid; ++j) e = s_step[e * nclass + s_class[p.in[base + j]]];
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

// Exclusive scan of the tile endomorphisms by one workgroup of 1024 lanes,
// restarting at the first tile of every input: each lane folds a contiguous
// run into (restarts inside, endomorphism since the last restart), a wave
// scan and a sequential pass over the 16 wave totals join the runs with
// (f1, a) . (f2, b) = (f1 | f2, f2 ? b : a then b), then each lane rewrites
// its run.
__global__ __launch_bounds__(1024) void lex_tile_scan(LexParams p) {
  __shared__ uint32_t s_wave[16], s_wflag[16];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const uint16_t* __restrict__ compose = p.compose;
  const uint32_t nendo = p.nendo;
  auto comp = [&](uint32_t a, uint32_t b) -> uint32_t { return compose[a * nendo + b]; };
  auto first = [&](uint32_t t) -> bool { return p.tile_first[p.tile_seg[t]] == t; };
  const uint32_t nt = p.ntiles;
  const uint32_t per = (nt + 1023) / 1024;
  const uint32_t lo = tid * per < nt ? tid * per : nt, hi = lo + per < nt ? lo + per : nt;
  uint32_t run = 0, rf = 0;
  for (uint32_t t = lo; t < hi; ++t) {
    if (first(t)) {
      run = 0;
      rf = 1;
    }
    run = comp(run, p.tile_endo[t]);
  }
  uint32_t inc = run, f = rf;
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    const uint32_t o = __shfl_up(inc, d, WAVE), of = __shfl_up(f, d, WAVE);
    if (lane >= (uint32_t)d) {
      inc = f ? inc : comp(o, inc);
      f |= of;
    }
  }
  if (lane == 63) {
    s_wave[wid] = inc;
    s_wflag[wid] = f;
  }
  __syncthreads();
  uint32_t wpre = 0;
  for (uint32_t i = 0; i < wid; ++i) wpre = s_wflag[i] ? s_wave[i] : comp(wpre, s_wave[i]);
  const uint32_t prev = __shfl_up(inc, 1, WAVE), pf = __shfl_up(f, 1, WAVE);
  uint32_t acc = lane == 0 ? wpre : (pf ? prev : comp(wpre, prev));
  for (uint32_t t = lo; t < hi; ++t) {
    if (first(t)) acc = 0;
    p.tile_prefix[t] = (uint16_t)acc;
    acc = comp(acc, p.tile_endo[t]);
  }
}

// ------------------------------------------------------------------- L3

constexpr uint32_t STAGE = 2048;  // kept tokens staged in LDS per round

// DFA state of the lane's first byte.
__device__ __forceinline__ uint32_t lane_state(const LexParams& p, uint32_t tile, uint32_t tid) {
  const uint32_t pre = p.compose[(uint32_t)p.tile_prefix[tile] * p.nendo + p.lane_prefix[(uint64_t)tile * WG + tid]];
  return p.endo_init[pre];
}

__global__ __launch_bounds__(WG) void lex_emit(LexParams p) {
  extern __shared__ __align__(16) uint8_t lds[];
  uint16_t* s_delta = reinterpret_cast<uint16_t*>(lds);  // nstate*256
  const uint32_t delta_bytes = ((uint32_t)p.nstate * 512u + 15u) & ~15u;
  uint32_t* s_start = reinterpret_cast<uint32_t*>(lds + delta_bytes);  // [STAGE] token start
  uint16_t* s_end = reinterpret_cast<uint16_t*>(s_start + STAGE);       // [STAGE] end, tile-relative
  uint8_t* s_term = reinterpret_cast<uint8_t*>(s_end + STAGE);          // [STAGE] terminal
  __shared__ uint8_t s_token[256];                                       // terminal each state accepts
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_wave_cnt[NW], s_wave_last[NW];
  __shared__ uint32_t s_pre_cnt, s_pre_last;
  const uint32_t tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
  for (int i = tid; i < p.nstate * 256; i += WG) s_delta[i] = p.delta[i];
  for (int i = tid; i < p.nstate; i += WG) s_token[i] = p.token[i];
  const int ignore = p.ignore;

  for (;;) {
    if (tid == 0) s_tile = atomicAdd(p.ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= p.ntiles) break;
    const TileSpan ts = tile_span(p, tile);
    const uint64_t tbase = ts.tbase, n = ts.seg_hi;
    const uint64_t base = tbase + (uint64_t)tid * IPT;
    const uint32_t valid = base >= n ? 0u : (base + IPT <= n ? (uint32_t)IPT : (uint32_t)(n - base));
    const bool full = ts.full;
    uint32_t a[IPT / 4 + 1];
    if (full) load_lane(p, base, ts.shift, a);
    const uint32_t s0 = lane_state(p, tile, tid);

    // walk 1: kept tokens closed in the lane and its last boundary.  Only
    // an input's last tile is partial: full tiles walk the registers fully
    // unrolled with no bounds checks, a partial one reads its bytes one by
    // one.
    // boundaries as bits (shift counts are inline constants, positions
    // above 64 would not be)
    uint32_t s = s0, cnt = 0, bm[IPT / 32] = {0, 0, 0, 0};
    auto step1 = [&](uint32_t byte, uint32_t j) {
      const uint32_t x = s_delta[s * 256 + byte];
      bm[j >> 5] |= ((x >> 8) & 1u) << (j & 31);
      cnt += (x >> 9) & 1u;
      s = x & 0xFFu;
    };
    if (full) {
#pragma unroll
      for (int k = 0; k < IPT / 4; ++k) {
        const uint32_t wk = lane_word(a, k, ts.shift, s);
#pragma unroll
        for (int b = 0; b < 4; ++b) step1((wk >> (8 * b)) & 0xFFu, 4 * k + b);
      }
    } else {
      for (uint32_t j = 0; j < valid; ++j) step1(p.in[base + j], j);
    }
    const bool owns_end = valid > 0 && base + valid == n;
    uint32_t fin_t = 0;
    bool fin_kept = false;
    if (owns_end) {
      fin_t = s_token[s];
      if ((int)fin_t == p.dead) {
        atomicOr(p.seg_flags + ts.seg, LEX_REJECT_FINAL);
      } else if ((int)fin_t != ignore) {
        fin_kept = true;
        ++cnt;
      }
    }
    uint32_t last = 0;
#pragma unroll
    for (int q = 0; q < IPT / 32; ++q)
      if (bm[q]) last = (uint32_t)base + 32 * q + 31 - __builtin_clz(bm[q]);

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
    const uint32_t pc = __shfl_up(icnt, 1, WAVE), pl = __shfl_up(ilast, 1, WAVE);
    const uint32_t ecnt = wc + (lane ? pc : 0u);
    const uint32_t elast = lane ? (wl > pl ? wl : pl) : wl;

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
          const int64_t idx = top - (int64_t)lane;
          const uint64_t st = idx >= 0 ? observe64(p.count_status + idx) : ST64_PRE;
          const uint64_t fl = st & ST64_MASK;
          const uint64_t zero = __ballot(fl == 0);
          const uint64_t prem = __ballot(fl == ST64_PRE);
          const int first_pre = prem ? __builtin_ctzll(prem) : WAVE;
          const int first_zero = zero ? __builtin_ctzll(zero) : WAVE;
          if (first_zero < first_pre) {  // a nearer tile has not published yet
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          uint32_t c = (int)lane <= first_pre && idx >= 0 ? (uint32_t)(st & 0x7FFFFFFFu) : 0u;
          uint32_t l = (int)lane <= first_pre && idx >= 0 ? (uint32_t)((st >> 31) & 0x7FFFFFFFu) : 0u;
#pragma unroll
          for (int d = 32; d >= 1; d >>= 1) {
            c += __shfl_xor(c, d, WAVE);
            const uint32_t o = __shfl_xor(l, d, WAVE);
            l = l > o ? l : o;
          }
          xc += c;
          xl = xl > l ? xl : l;
          if (first_pre < WAVE) break;
          top -= WAVE;
        }
        if (lane == 0) {
          const uint32_t ic = xc + tc, il = xl > tl ? xl : tl;
          publish64(p.count_status + tile, ST64_PRE | (uint64_t)ic | ((uint64_t)il << 31));
        }
      }
      if (lane == 0) {
        s_pre_cnt = xc;
        s_pre_last = xl;
      }
    }
    __syncthreads();
    const uint64_t slot0 = s_pre_cnt;  // output slot of the tile's first kept token
    uint32_t start0 = s_pre_last > elast ? s_pre_last : elast;
    start0 = start0 > (uint32_t)ts.seg_lo ? start0 : (uint32_t)ts.seg_lo;  // an input's first token starts at its first byte
    if (tid == 0 && ts.first) p.seg_tok[ts.seg] = slot0;
    if (tid == 0 && tile == p.ntiles - 1) p.seg_tok[p.nseg] = slot0 + tc;
    bool overflow = false;
    const uint32_t lane_off = tid * IPT;

    // walk 2, per round of STAGE slots: the lanes owning slots of the round
    // re-walk and stage (terminal, start, end); then the workgroup writes the
    // round out with consecutive lanes on consecutive slots
    for (uint32_t lo = 0; lo < tc; lo += STAGE) {
      const uint32_t hi = lo + STAGE;
      if (cnt > 0 && ecnt < hi && ecnt + cnt > lo) {
        // the lane's offsets are opaque here, so the 128 positions are not
        // kept live in registers across the rounds
        uint32_t vbase = (uint32_t)base, voff = lane_off;
        asm volatile("" : "+v"(vbase), "+v"(voff));
        uint32_t st = s0, k = ecnt, start = start0;
        auto step2 = [&](uint32_t byte, uint32_t j) {
          const uint32_t x = s_delta[st * 256 + byte];
          if (x & 0x200u) {  // a kept token closes before byte j
            const uint32_t i = k - lo;
            if (i < STAGE) {
              s_term[i] = s_token[st];
              s_start[i] = start;
              s_end[i] = (uint16_t)(voff + j);
            }
            ++k;
          }
          start = (x & 0x100u) ? vbase + j : start;
          st = x & 0xFFu;
        };
        if (full) {
#pragma unroll
          for (int k = 0; k < IPT / 4; ++k) {
            const uint32_t wk = lane_word(a, k, ts.shift, st);
#pragma unroll
            for (int b = 0; b < 4; ++b) step2((wk >> (8 * b)) & 0xFFu, 4 * k + b);
          }
        } else {
          for (uint32_t j = 0; j < valid; ++j) step2(p.in[base + j], j);
        }
        if (fin_kept) {
          const uint32_t i = k - lo;
          if (i < STAGE) {
            s_term[i] = (uint8_t)fin_t;
            s_start[i] = start;
            s_end[i] = (uint16_t)(lane_off + valid);
          }
        }
      }
      __syncthreads();
      const uint32_t m = tc - lo < STAGE ? tc - lo : STAGE;
      const uint64_t max_len = p.max_len;
      for (uint32_t i = tid; i < m; i += WG) {
        const uint64_t slot = slot0 + lo + i;
        const uint32_t b = s_start[i];
        const uint64_t len = tbase + s_end[i] - b;
        overflow |= len > max_len;
        p.terminals[slot] = s_term[i];
        p.starts[slot] = (int64_t)(b - ts.seg_lo);
        store_length(p, slot, len);
      }
      __syncthreads();
    }
    if (overflow) atomicOr(p.seg_flags + ts.seg, LEX_REJECT_OVERFLOW);
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
size_t lex_emit_lds(const LexParams& p) {
  return (((size_t)p.nstate * 512 + 15) & ~size_t(15)) + (size_t)STAGE * (4 + 2 + 1);
}
__global__ void tile_map(const uint32_t* tile_first, uint32_t nseg, uint32_t ntiles, uint32_t* tile_seg) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  uint32_t lo = 0, hi = nseg;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (tile_first[mid] <= t)
      lo = mid;
    else
      hi = mid;
  }
  tile_seg[t] = lo;
}


hipError_t lex_grid_size(const LexParams& p, int* fold_grid, int* emit_grid) {
  hipError_t e = occupancy_grid(lex_fold, lex_fold_lds(p), fold_grid);
  if (e != hipSuccess) return e;
  return occupancy_grid(lex_emit, lex_emit_lds(p), emit_grid);
}

hipError_t lex_launch(const LexParams& p, int fold_grid, int emit_grid, hipStream_t st, KernelTimer& tm) {
  const uint32_t nt = p.ntiles;
  int g1 = fold_grid < (int)nt ? fold_grid : (int)nt;
  int g3 = emit_grid < (int)nt ? emit_grid : (int)nt;
  hipError_t e;
  hipLaunchKernelGGL(tile_map, dim3((nt + 255) / 256), dim3(256), 0, st, p.tile_first, p.nseg, nt, p.tile_seg);
  if ((e = tm.begin(st, "lex_fold")) != hipSuccess) return e;
  hipLaunchKernelGGL(lex_fold, dim3(g1), dim3(WG), lex_fold_lds(p), st, p);
  if ((e = tm.end(st)) != hipSuccess || (e = tm.begin(st, "lex_tile_scan")) != hipSuccess) return e;
  hipLaunchKernelGGL(lex_tile_scan, dim3(1), dim3(1024), 0, st, p);
  if ((e = tm.end(st)) != hipSuccess || (e = tm.begin(st, "lex_emit")) != hipSuccess) return e;
  hipLaunchKernelGGL(lex_emit, dim3(g3), dim3(WG), lex_emit_lds(p), st, p);
  if ((e = tm.end(st)) != hipSuccess) return e;
  return hipGetLastError();
}

}  // namespace alpacc
