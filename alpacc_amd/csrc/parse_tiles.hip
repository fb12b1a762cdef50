// Tiled LLP parser for gfx950: three passes over tiles of 1024 framed
// positions instead of one pass per intermediate array.
//
// Semantics are those of backends/c/parser.c: parse_test (:49-98) checks that
// every window has a pair and that the concatenated bracket strings are
// balanced and symbol-matched; compute_parents (:103-118) and compact_tree
// (:126-153) build the productions-only tree.  Both checks are bracket
// matching problems:
//   * the stack-change strings are brackets directly;
//   * every derivation slot pops one open child slot of its parent and pushes
//     `arity` new ones, so a slot's parent is the owner of the open slot its
//     pop matches (the root pops a virtual slot owned by index 0).
// Depth moves by one per bracket, so the match of a pop with depth-before D
// is the last earlier item whose depth is below D.
//
//   A  pair_profile  window -> pair id per position; per tile: bracket and
//                    derivation lengths, terminal count and the (sum, min)
//                    depth profiles of both bracket streams;
//   S  tile_scan     prefixes over tiles: offsets, starting depths, absolute
//                    tile minima (for the far search) and the offsets of the
//                    per-tile lists of open pushes and unmatched pops;
//   B  tile_resolve  per tile in LDS: expand brackets and slots, match
//                    locally, write the tree; pushes still open at the tile
//                    end go to the tile's open list, pops without a local
//                    match to its unmatched list;
//   C  far_resolve   each unmatched pop finds the last earlier tile whose
//                    minimum is below its depth (a 32-ary min tree over
//                    tiles) and reads the open push at its level there.
// Traffic per token: tokens read twice, a u16 pair id written and read, and
// the outputs (JSON: about 22 bytes per token in total).
#include <hip/hip_runtime.h>

#include "../../include/alpacc_hip.h"
#include "device.hpp"
#include "parser.hpp"

namespace alpacc {

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return fail(ALPACC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

namespace {

constexpr int TWG = 256;         // lanes per workgroup of pass A
constexpr int TPL = 4;           // positions per lane of pass A
constexpr int TPOS = TWG * TPL;  // positions per tile

constexpr uint32_t BAD_PAIR = 1u, BAD_DEPTH = 2u, BAD_MATCH = 4u, BAD_COUNT = 8u, BAD_ROOT = 16u;
constexpr int64_t TOKEN_TARGET = 1ll << 62;

struct Pairs {
  const uint8_t* tok;
  uint64_t m;
  int q, k, base, empty_t, start_t, end_t;
  const uint16_t* window;
  const uint2* head;  // first 8 bytes of each 32-byte pair record
};

__device__ __forceinline__ uint32_t pair_of(const Pairs& a, uint64_t i) {
  uint32_t code = 0;
  const int w = a.q + a.k;
  for (int j = 0; j < w; ++j) {
    int64_t f = (int64_t)i - a.q + j;
    uint32_t t;
    if (f < 0 || f >= (int64_t)a.m)
      t = a.empty_t;
    else if (f == 0)
      t = a.start_t;
    else if (f == (int64_t)a.m - 1)
      t = a.end_t;
    else
      t = a.tok[f - 1];
    if (t >= (uint32_t)a.base) return 0xFFFFu;
    code = code * a.base + t;
  }
  return a.window[code];
}

// (sum, min-prefix) monoid: a then b.
struct SM {
  int32_t s, m;
};
__device__ __forceinline__ SM sm_cat(SM a, SM b) { return {a.s + b.s, a.m < a.s + b.m ? a.m : a.s + b.m}; }

struct TileRec {
  int32_t nb, np, nt, bs, bm, ss, sm, pad;
};

struct TileInfo {
  int64_t b_off, p_off, t_off;
  int32_t bstart, sstart, bmin, smin;
  int32_t ob_off, ub_off, os_off, us_off;
};

// Ordered inclusive scan of SM over the block (lane order = position order).
__device__ __forceinline__ SM wave_scan_sm(SM v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    SM o{__shfl_up(v.s, d, WAVE), __shfl_up(v.m, d, WAVE)};
    if (lane >= (uint32_t)d) v = sm_cat(o, v);
  }
  return v;
}

// ---------------------------------------------------------------- pass A

__global__ __launch_bounds__(TWG) void pair_profile(Pairs a, uint16_t* pid, TileRec* rec, uint32_t* flags) {
  const uint64_t i0 = (uint64_t)blockIdx.x * TPOS + (uint64_t)threadIdx.x * TPL;
  int32_t nb = 0, np = 0, nt = 0;
  SM b{0, 0}, s{0, 0};
  bool bad = false;
  uint16_t ids[TPL];
#pragma unroll
  for (int r = 0; r < TPL; ++r) {
    uint64_t i = i0 + r;
    uint32_t id = 0xFFFFu;
    if (i < a.m) {
      id = pair_of(a, i);
      if (id == 0xFFFFu) {
        bad = true;
      } else {
        // [brackets, productions, terminals, bracket sum, bracket min, slot sum, slot min, 0]
        const uint2 h = a.head[4 * id];
        nb += h.x & 0xFFu;
        np += (h.x >> 8) & 0xFFu;
        nt += (h.x >> 16) & 0xFFu;
        b = sm_cat(b, SM{(int8_t)(h.x >> 24), (int8_t)h.y});
        s = sm_cat(s, SM{(int8_t)(h.y >> 8), (int8_t)(h.y >> 16)});
      }
    }
    ids[r] = (uint16_t)id;
  }
  if (bad) atomicOr(flags, BAD_PAIR);
  if (i0 + TPL <= a.m) {
    uint2 v;
    v.x = ids[0] | ((uint32_t)ids[1] << 16);
    v.y = ids[2] | ((uint32_t)ids[3] << 16);
    *reinterpret_cast<uint2*>(pid + i0) = v;
  } else {
    for (int r = 0; r < TPL; ++r)
      if (i0 + r < a.m) pid[i0 + r] = ids[r];
  }
  // block reduction (ordered for the min-prefix monoids)
  __shared__ int32_t wc[3][TWG / WAVE];
  __shared__ SM wb[TWG / WAVE], ws[TWG / WAVE];
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t cb = wave_scan_add32((uint32_t)nb), cp = wave_scan_add32((uint32_t)np), ct = wave_scan_add32((uint32_t)nt);
  SM ib = wave_scan_sm(b), is = wave_scan_sm(s);
  if (lane == 63) {
    wc[0][wid] = (int32_t)cb;
    wc[1][wid] = (int32_t)cp;
    wc[2][wid] = (int32_t)ct;
    wb[wid] = ib;
    ws[wid] = is;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    TileRec r{0, 0, 0, 0, 0, 0, 0, 0};
    SM tb{0, 0}, tsm{0, 0};
    for (int w = 0; w < TWG / WAVE; ++w) {
      r.nb += wc[0][w];
      r.np += wc[1][w];
      r.nt += wc[2][w];
      tb = sm_cat(tb, wb[w]);
      tsm = sm_cat(tsm, ws[w]);
    }
    r.bs = tb.s;
    r.bm = tb.m;
    r.ss = tsm.s;
    r.sm = tsm.m;
    rec[blockIdx.x] = r;
  }
}

// ---------------------------------------------------------------- pass S

struct ScanOut {
  int64_t nb, np, nt;  // totals
  int64_t bend, send;  // final depths
};

// The nine running sums over tiles: brackets, slots, terminals, bracket
// depth, slot depth, open pushes / unmatched pops of both streams.
constexpr int NV = 9;
constexpr int SB = 1024;  // tiles per scan block

__device__ __forceinline__ void tile_terms(const TileRec& r, int64_t* v) {
  v[0] = r.nb;
  v[1] = r.np;
  v[2] = r.nt;
  v[3] = r.bs;
  v[4] = r.ss;
  v[5] = r.bs - r.bm;
  v[6] = -r.bm;
  v[7] = r.ss - r.sm;
  v[8] = -r.sm;
}

// Block-wide exclusive scan of NV int64 values (256 lanes); totals out.
__device__ __forceinline__ void block_scan_nv(int64_t* v, int64_t* tot) {
  __shared__ int64_t ws[NV][4];
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    int64_t inc = (int64_t)wave_scan_add64((uint64_t)v[k]);
    if (lane == 63) ws[k][wid] = inc;
    v[k] = inc - v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    int64_t o = 0, t = 0;
    for (uint32_t w = 0; w < 4; ++w) {
      if (w < wid) o += ws[k][w];
      t += ws[k][w];
    }
    v[k] += o;
    tot[k] = t;
  }
  __syncthreads();
}

// S1: per block of SB tiles, the sums.
__global__ __launch_bounds__(256) void tile_block_sums(const TileRec* rec, uint32_t ntiles, int64_t* bsum) {
  int64_t v[NV] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const uint32_t t0 = blockIdx.x * SB + threadIdx.x * (SB / 256);
  for (int i = 0; i < SB / 256; ++i) {
    if (t0 + i >= ntiles) break;
    int64_t x[NV];
    tile_terms(rec[t0 + i], x);
    for (int k = 0; k < NV; ++k) v[k] += x[k];
  }
  int64_t tot[NV];
  block_scan_nv(v, tot);
  if (threadIdx.x == 0)
    for (int k = 0; k < NV; ++k) bsum[(uint64_t)blockIdx.x * NV + k] = tot[k];
}

// S2: exclusive scan of the block sums (one workgroup), totals and checks.
__global__ __launch_bounds__(256) void tile_block_scan(int64_t* bsum, uint32_t nblk, ScanOut* out, uint64_t ntokens,
                                                       uint32_t* flags) {
  __shared__ int64_t carry[NV];
  if (threadIdx.x < NV) carry[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t b0 = 0; b0 < nblk; b0 += 256) {
    const uint32_t b = b0 + threadIdx.x;
    int64_t v[NV];
    for (int k = 0; k < NV; ++k) v[k] = b < nblk ? bsum[(uint64_t)b * NV + k] : 0;
    int64_t tot[NV];
    block_scan_nv(v, tot);
    if (b < nblk)
      for (int k = 0; k < NV; ++k) bsum[(uint64_t)b * NV + k] = carry[k] + v[k];
    __syncthreads();
    if (threadIdx.x < NV) carry[threadIdx.x] += tot[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    ScanOut o{carry[0], carry[1], carry[2], carry[3], carry[4]};
    *out = o;
    if (carry[3] != 0) atomicOr(flags, BAD_DEPTH);
    if (ntokens != ~0ull && (uint64_t)carry[2] != ntokens) atomicOr(flags, BAD_COUNT);
  }
}

// S3: every tile's prefixes, minima and list offsets.
__global__ __launch_bounds__(256) void tile_downsweep(const TileRec* rec, uint32_t ntiles, const int64_t* bsum,
                                                      TileInfo* info, int32_t* mb, int32_t* ms, uint32_t* flags) {
  int64_t v[NV] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const uint32_t t0 = blockIdx.x * SB + threadIdx.x * (SB / 256);
  for (int i = 0; i < SB / 256; ++i) {
    if (t0 + i >= ntiles) break;
    int64_t x[NV];
    tile_terms(rec[t0 + i], x);
    for (int k = 0; k < NV; ++k) v[k] += x[k];
  }
  int64_t tot[NV];
  block_scan_nv(v, tot);
  for (int k = 0; k < NV; ++k) v[k] += bsum[(uint64_t)blockIdx.x * NV + k];
  bool neg = false;
  for (int i = 0; i < SB / 256; ++i) {
    const uint32_t t = t0 + i;
    if (t >= ntiles) break;
    const TileRec r = rec[t];
    TileInfo ti;
    ti.b_off = v[0];
    ti.p_off = v[1];
    ti.t_off = v[2];
    ti.bstart = (int32_t)v[3];
    ti.sstart = (int32_t)(v[4] + 1);  // the virtual root slot
    ti.bmin = r.bm;
    ti.smin = r.sm;
    ti.ob_off = (int32_t)v[5];
    ti.ub_off = (int32_t)v[6];
    ti.os_off = (int32_t)v[7];
    ti.us_off = (int32_t)v[8];
    info[t] = ti;
    mb[t] = ti.bstart + r.bm;
    ms[t] = ti.sstart + r.sm;
    neg |= ti.bstart + r.bm < 0;
    int64_t x[NV];
    tile_terms(r, x);
    for (int k = 0; k < NV; ++k) v[k] += x[k];
  }
  if (neg) atomicOr(flags, BAD_DEPTH);
}

// ---------------------------------------------------------------- pass B

struct Resolve {
  const uint16_t* pid;
  uint64_t m;
  const TileInfo* info;
  const uint4* rec;          // 32-byte pair records
  const uint8_t* prod_info;  // arity | terminal << 7
  int nprod;
  int bcap, scap;     // per-lane LDS stack room (brackets, slots)
  uint16_t* ob_list;  // open pushes (bracket symbols)
  uint16_t* ub_list;  // unmatched pops (bracket symbols)
  int64_t* os_list;   // open child slots (owner's compact index)
  int64_t* us_list;   // unmatched slots (output target)
  uint8_t* tree_prod;
  int64_t* tree_par;
  int64_t* tok_par;
  uint8_t* deriv;  // flat derivation (parser-only mode) or null
  int tree;        // compute parents and the compact tree
  uint32_t* flags;
};

constexpr int RWG = 256;         // lanes per tile in pass B
constexpr int RPL = TPOS / RWG;  // positions per lane
constexpr int RLV = 8;           // sparse-table levels: blocks of 1..128 lanes
constexpr int RNW = RWG / WAVE;
constexpr uint16_t TERM_TARGET = 0x8000u;

// Bracket j (0..7) and production j of a pair record held as two 16-byte words.
__device__ __forceinline__ uint32_t rec_bracket(const uint4& h, const uint4& t, int j) {
  const uint64_t w = j < 4 ? (((uint64_t)h.w << 32) | h.z) : (((uint64_t)t.y << 32) | t.x);
  return (uint32_t)(w >> (16 * (j & 3))) & 0xFFFFu;
}
__device__ __forceinline__ uint32_t rec_prod(const uint4& t, int j) {
  const uint64_t w = ((uint64_t)t.w << 32) | t.z;
  return (uint32_t)(w >> (8 * j)) & 0xFFu;
}

// Last lane l <= from whose minimum (tb[0][l]) is below D, or a negative
// value: tb[k][l] is the minimum over lanes [l - 2^k + 1, l] (clipped at 0),
// and maximal blocks with minimum >= D are skipped, largest first.
__device__ __forceinline__ int prev_below(const int16_t (*tb)[RWG], int from, int32_t D) {
  int l = from;
  if (l < 0 || tb[0][l] < D) return l;
#pragma unroll
  for (int k = RLV - 1; k >= 0; --k)
    if (l >= 0 && tb[k][l] >= D) l -= 1 << k;
  return l;
}

// Exclusive minimum over the later lanes of the workgroup.
__device__ __forceinline__ int32_t suffix_min_after(int32_t v, int32_t* wtot) {
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int32_t inc = v;
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    const int32_t o = __shfl_down(inc, d, WAVE);
    if (lane + d < (uint32_t)WAVE) inc = o < inc ? o : inc;
  }
  if (lane == 0) wtot[wid] = inc;
  __syncthreads();
  int32_t ex = __shfl_down(inc, 1, WAVE);
  if (lane == WAVE - 1) ex = 0x7FFFFFFF;
  for (uint32_t w = wid + 1; w < (uint32_t)RNW; ++w) ex = wtot[w] < ex ? wtot[w] : ex;
  return ex;
}

// Every lane runs the reference's stack algorithms over its RPL positions:
// brackets on a symbol stack, derivation slots on a stack of open child slots
// (one entry per slot still owed, holding the owner's tile-local compact
// index).  The stacks live in LDS interleaved by lane (entry e of lane t at
// e * RWG + t, so the lanes of a wave touch distinct banks); pops that find
// the lane's stack empty are recorded from the top entry down.
// A lane's open entries sit at consecutive depth levels just above the
// lane's minimum, so a pop left over by a later lane at tile-relative level D
// reads entry D - min - 1 of the last earlier lane whose minimum is below D.
// Depths are tile-relative and fit int16 (the host checks the bound).
__global__ __launch_bounds__(RWG) void tile_resolve(Resolve r) {
  extern __shared__ __align__(16) uint16_t stk[];  // [bcap][RWG] brackets, then [scap][RWG] slots
  __shared__ int16_t tbb[RLV][RWG], tbs[RLV][RWG];
  __shared__ uint8_t s_pinfo[256];
  __shared__ int32_t ws[4][RNW];
  __shared__ int32_t wmin[2][RNW];
  const uint32_t tile = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  for (int i = tid; i < r.nprod; i += RWG) s_pinfo[i] = r.prod_info[i];
  const TileInfo ti = r.info[tile];
  const uint64_t i0 = (uint64_t)tile * TPOS + (uint64_t)tid * RPL;
  const int bcap = r.bcap, scap = r.scap;

  // phase 0: pair records, lane offsets (slots, terminals) and starting depths
  uint32_t ids[RPL];
  if (i0 + RPL <= r.m) {
    const uint2 v = *reinterpret_cast<const uint2*>(r.pid + i0);
    ids[0] = v.x & 0xFFFFu;
    ids[1] = v.x >> 16;
    ids[2] = v.y & 0xFFFFu;
    ids[3] = v.y >> 16;
  } else {
#pragma unroll
    for (int k = 0; k < RPL; ++k) ids[k] = i0 + k < r.m ? r.pid[i0 + k] : 0xFFFFu;
  }
  uint4 rh[RPL], rt[RPL];
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    const uint32_t id = ids[k] == 0xFFFFu ? 0u : ids[k];
    rh[k] = r.rec[2 * id];
    rt[k] = r.rec[2 * id + 1];
  }
  int32_t np = 0, nt = 0, bd = 0, sd = 0;
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    if (ids[k] == 0xFFFFu) continue;
    np += (rh[k].x >> 8) & 0xFFu;
    nt += (rh[k].x >> 16) & 0xFFu;
    bd += (int8_t)(rh[k].x >> 24);
    sd += (int8_t)(rh[k].y >> 8);
  }
  int32_t x0 = (int32_t)wave_scan_add32((uint32_t)np), x1 = (int32_t)wave_scan_add32((uint32_t)nt),
          x2 = (int32_t)wave_scan_add32((uint32_t)bd), x3 = (int32_t)wave_scan_add32((uint32_t)sd);
  if (lane == 63) {
    ws[0][wid] = x0;
    ws[1][wid] = x1;
    ws[2][wid] = x2;
    ws[3][wid] = x3;
  }
  __syncthreads();
  int32_t slot = x0 - np, tinc = x1 - nt, bS = x2 - bd, sS = x3 - sd;
  for (uint32_t w = 0; w < wid; ++w) {
    slot += ws[0][w];
    tinc += ws[1][w];
    bS += ws[2][w];
    sS += ws[3][w];
  }

  // phase 1: stacks
  const int64_t p_off = ti.p_off, t_off = ti.t_off, c_off = p_off - t_off;
  uint16_t* const bst = stk + tid;
  uint16_t* const sst = stk + RWG * bcap + tid;
  int bsp = 0, bun = 0, ssp = 0, sun = 0;
  int32_t bdep = 0, bmin = 0, sdep = 0, smin = 0;
  bool bad = false;
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    if (ids[k] == 0xFFFFu) continue;
    const uint4 h = rh[k], t = rt[k];
    const int sl = h.x & 0xFFu, pl = (h.x >> 8) & 0xFFu;
    for (int j = 0; j < sl; ++j) {
      const uint32_t b = rec_bracket(h, t, j);
      if (b & 0x8000u) {
        bst[RWG * bsp++] = (uint16_t)b;
        ++bdep;
      } else {
        if (bsp > 0) {
          bad |= bst[RWG * --bsp] != (uint16_t)(b | 0x8000u);
        } else {
          bst[RWG * (bcap - 1 - bun)] = (uint16_t)b;
          ++bun;
        }
        --bdep;
        bmin = bdep < bmin ? bdep : bmin;
      }
    }
    for (int j = 0; j < pl; ++j) {
      const uint32_t pr = rec_prod(t, j);
      const int64_t J = p_off + slot;
      if (r.deriv) r.deriv[J] = (uint8_t)pr;
      if (r.tree) {
        const uint32_t info = s_pinfo[pr];
        const bool term = info & 0x80u;
        tinc += term;
        const int32_t self_c = slot - tinc;  // tile-local compact index (productions)
        if (!term) r.tree_prod[c_off + self_c] = (uint8_t)pr;
        if (ssp > 0) {
          const int64_t pc = c_off + sst[RWG * --ssp];
          if (term)
            r.tok_par[t_off + tinc - 1] = pc;
          else
            r.tree_par[c_off + self_c] = pc;
        } else {
          sst[RWG * (scap - 1 - sun)] = term ? (uint16_t)(TERM_TARGET | (tinc - 1)) : (uint16_t)self_c;
          ++sun;
        }
        --sdep;
        smin = sdep < smin ? sdep : smin;
        const int a = info & 0x7Fu;
        for (int q = 0; q < a; ++q) sst[RWG * ssp++] = (uint16_t)self_c;
        sdep += a;
      }
      ++slot;
    }
  }
  if (bad) atomicOr(r.flags, BAD_MATCH);
  const int32_t lbm = bS + bmin, lsm = sS + smin;
  tbb[0][tid] = (int16_t)lbm;
  tbs[0][tid] = (int16_t)lsm;
  const int32_t bsuf = suffix_min_after(lbm, wmin[0]);
  const int32_t ssuf = suffix_min_after(lsm, wmin[1]);
  // level 0 is visible: suffix_min_after synchronised after it was written
#pragma unroll
  for (int k = 1; k < RLV; ++k) {
    const int h = 1 << (k - 1);
    int16_t a = tbb[k - 1][tid], c = tbs[k - 1][tid];
    if ((int)tid >= h) {
      const int16_t a2 = tbb[k - 1][tid - h], c2 = tbs[k - 1][tid - h];
      a = a2 < a ? a2 : a;
      c = c2 < c ? c2 : c;
    }
    tbb[k][tid] = a;
    tbs[k][tid] = c;
    __syncthreads();
  }

  // phase 2: pops left over by this lane (depth falling, so the matching
  // lane only moves back); entries left open at the tile end
  int l = (int)tid - 1;
  for (int u = 0; u < bun; ++u) {
    const int32_t D = bS - u;
    const uint32_t b = bst[RWG * (bcap - 1 - u)];
    l = prev_below(tbb, l, D);
    if (l >= 0) {
      const uint16_t push = stk[RWG * (D - tbb[0][l] - 1) + l];
      bad |= push != (uint16_t)(b | 0x8000u);
    } else {
      r.ub_list[ti.ub_off + (-D)] = (uint16_t)b;
    }
  }
  for (int e = 0; e < bsp; ++e) {
    const int32_t v = lbm + 1 + e;
    if (v <= bsuf) r.ob_list[ti.ob_off + (v - ti.bmin - 1)] = bst[RWG * e];
  }
  if (r.tree) {
    l = (int)tid - 1;
    for (int u = 0; u < sun; ++u) {
      const int32_t D = sS - u;
      const uint32_t tg = sst[RWG * (scap - 1 - u)];
      l = prev_below(tbs, l, D);
      const bool term = tg & TERM_TARGET;
      if (l >= 0) {
        const int64_t pc = c_off + stk[RWG * (bcap + D - tbs[0][l] - 1) + l];
        if (term)
          r.tok_par[t_off + (tg & 0x7FFFu)] = pc;
        else
          r.tree_par[c_off + tg] = pc;
      } else {
        r.us_list[ti.us_off + (-D)] = term ? (TOKEN_TARGET | (t_off + (tg & 0x7FFFu))) : (c_off + tg);
      }
    }
    for (int e = 0; e < ssp; ++e) {
      const int32_t v = lsm + 1 + e;
      if (v <= ssuf) r.os_list[ti.os_off + (v - ti.smin - 1)] = c_off + sst[RWG * e];
    }
  }
  if (bad) atomicOr(r.flags, BAD_MATCH);
}

// ---------------------------------------------------------------- pass C

constexpr int MT_LEVELS = 8;
struct TileTree {
  const int32_t* lv[MT_LEVELS];
  uint64_t n[MT_LEVELS];
  int levels;
};

// last tile t' < t with min < v, or -1
__device__ int64_t tile_prev_less(const TileTree& t, uint64_t i, int32_t v) {
  uint64_t idx = i;
  for (int L = 0; L < t.levels; ++L) {
    const int32_t* a = t.lv[L];
    uint64_t lo = idx & ~31ull;
    for (uint64_t j = idx; j > lo;) {
      --j;
      if (a[j] < v) {
        uint64_t k = j;
        for (int D = L; D > 0; --D) {
          const int32_t* b = t.lv[D - 1];
          uint64_t c0 = k * 32, c1 = c0 + 32 < t.n[D - 1] ? c0 + 32 : t.n[D - 1];
          uint64_t c = c1;
          while (c > c0) {
            --c;
            if (b[c] < v) break;
          }
          k = c;
        }
        return (int64_t)k;
      }
    }
    idx >>= 5;
  }
  return -1;
}

struct Far {
  const TileInfo* info;
  TileTree tb, ts;
  const uint16_t* ob_list;
  const uint16_t* ub_list;
  const int64_t* os_list;
  const int64_t* us_list;
  int64_t* tree_par;
  int64_t* tok_par;
  int tree;
  uint32_t* flags;
};

__global__ __launch_bounds__(TWG) void far_resolve(Far f) {
  const uint32_t tile = blockIdx.x;
  const TileInfo ti = f.info[tile];
  bool bad = false;
  for (int u = threadIdx.x; u < -ti.bmin; u += TWG) {
    int32_t D = ti.bstart - u;  // depth before the pop
    uint16_t b = f.ub_list[ti.ub_off + u];
    int64_t t2 = tile_prev_less(f.tb, tile, D);
    if (t2 < 0) {
      bad = true;
      continue;
    }
    const TileInfo& o = f.info[t2];
    uint16_t push = f.ob_list[o.ob_off + (D - o.bstart - o.bmin - 1)];
    bad |= push != (uint16_t)(b | 0x8000u);
  }
  for (int u = threadIdx.x; f.tree && u < -ti.smin; u += TWG) {
    int32_t D = ti.sstart - u;  // open child slots before the pop
    int64_t target = f.us_list[ti.us_off + u];
    int64_t t2 = tile_prev_less(f.ts, tile, D);
    int64_t pc;
    if (t2 < 0) {
      if (D != 1) bad = true;
      pc = 0;  // the root is its own parent
    } else {
      const TileInfo& o = f.info[t2];
      pc = f.os_list[o.os_off + (D - o.sstart - o.smin - 1)];
    }
    if (target & TOKEN_TARGET)
      f.tok_par[target & ~TOKEN_TARGET] = pc;
    else
      f.tree_par[target] = pc;
  }
  if (bad) atomicOr(f.flags, BAD_ROOT);
}

__global__ __launch_bounds__(TWG) void tiles_min_level(const int32_t* in, uint64_t n_in, int32_t* out, uint64_t n_out) {
  uint64_t c = (uint64_t)blockIdx.x * TWG + threadIdx.x;
  if (c >= n_out) return;
  uint64_t lo = c * 32, hi = lo + 32 < n_in ? lo + 32 : n_in;
  int32_t m = in[lo];
  for (uint64_t i = lo + 1; i < hi; ++i) m = in[i] < m ? in[i] : m;
  out[c] = m;
}

int build_tile_tree(const int32_t* lv0, uint64_t n, int32_t* scratch, hipStream_t st, TileTree* t) {
  t->levels = 1;
  t->lv[0] = lv0;
  t->n[0] = n;
  int32_t* p = scratch;
  uint64_t m = n;
  while (m > 32) {
    if (t->levels == MT_LEVELS) return fail(ALPACC_ERR_LIMIT, "tile tree too deep");
    uint64_t nn = (m + 31) / 32;
    hipLaunchKernelGGL(tiles_min_level, dim3((uint32_t)((nn + TWG - 1) / TWG)), dim3(TWG), 0, st,
                       t->lv[t->levels - 1], m, p, nn);
    t->lv[t->levels] = p;
    t->n[t->levels] = nn;
    t->levels++;
    p += nn;
    m = nn;
  }
  HIP_TRY(hipGetLastError());
  return 0;
}

uint64_t tree_scratch(uint64_t n) {
  uint64_t tot = 0;
  while (n > 32) {
    n = (n + 31) / 32;
    tot += n;
  }
  return tot + 1;
}

}  // namespace

int parse_on_device(Context& c, const uint8_t* d_tok, uint64_t n, bool tree, hipStream_t st, uint64_t* num_prods) {
  const GrammarDesc* g = c.g;
  *num_prods = 0;
  const int bcap = RPL * g->max_brackets, scap = RPL * g->max_slot_items;
  const size_t lds_b = (size_t)RWG * (bcap + scap) * 2;
  // tile-relative depths are int16 in pass B
  if ((int64_t)TPOS * g->max_brackets > 0x7FFF || (int64_t)TPOS * g->max_slot_items > 0x7FFF || lds_b > 96 * 1024)
    return fail(ALPACC_ERR_LIMIT, "grammar windows too long for the tile stacks");
  const uint64_t m = n + 2;
  const uint64_t ntiles = (m + TPOS - 1) / TPOS;
  if (ntiles > 0x7FFFFFFFull) return fail(ALPACC_ERR_LIMIT, "too many tiles");
  int r;
  // workspace: pid | TileRec | TileInfo | mb | ms | trees | block sums | ScanOut
  const uint64_t tt = tree_scratch(ntiles);
  const uint32_t nsb = (uint32_t)((ntiles + SB - 1) / SB);
  size_t off_pid = 0;
  size_t off_rec = (off_pid + m * 2 + 255) & ~size_t(255);
  size_t off_info = (off_rec + ntiles * sizeof(TileRec) + 255) & ~size_t(255);
  size_t off_mb = (off_info + ntiles * sizeof(TileInfo) + 255) & ~size_t(255);
  size_t off_ms = (off_mb + ntiles * 4 + 255) & ~size_t(255);
  size_t off_tb = (off_ms + ntiles * 4 + 255) & ~size_t(255);
  size_t off_ts = (off_tb + tt * 4 + 255) & ~size_t(255);
  size_t off_bs = (off_ts + tt * 4 + 255) & ~size_t(255);
  size_t off_out = (off_bs + (size_t)nsb * NV * 8 + 255) & ~size_t(255);
  size_t total = off_out + sizeof(ScanOut) + 256;
  if ((r = c.ps_a.reserve(total))) return r;
  uint8_t* base = c.ps_a.as<uint8_t>();
  uint16_t* pid = reinterpret_cast<uint16_t*>(base + off_pid);
  TileRec* rec = reinterpret_cast<TileRec*>(base + off_rec);
  TileInfo* info = reinterpret_cast<TileInfo*>(base + off_info);
  int32_t* mb = reinterpret_cast<int32_t*>(base + off_mb);
  int32_t* ms = reinterpret_cast<int32_t*>(base + off_ms);
  ScanOut* sout = reinterpret_cast<ScanOut*>(base + off_out);
  int64_t* bsum = reinterpret_cast<int64_t*>(base + off_bs);
  uint32_t* flags = reinterpret_cast<uint32_t*>(c.small.as<uint64_t>() + 2);
  if ((r = zero_words(flags, 4, st))) return r;

  if (c.timing) HIP_TRY(hipEventRecord(c.ev[2], st));
  Pairs a{d_tok, m, g->q, g->k, g->window_base, g->empty_t, g->start_t, g->end_t, c.d_window,
          reinterpret_cast<const uint2*>(c.d_pair_rec)};
  hipLaunchKernelGGL(pair_profile, dim3((uint32_t)ntiles), dim3(TWG), 0, st, a, pid, rec, flags);
  hipLaunchKernelGGL(tile_block_sums, dim3(nsb), dim3(256), 0, st, rec, (uint32_t)ntiles, bsum);
  hipLaunchKernelGGL(tile_block_scan, dim3(1), dim3(256), 0, st, bsum, nsb, sout, tree ? n : ~0ull, flags);
  hipLaunchKernelGGL(tile_downsweep, dim3(nsb), dim3(256), 0, st, rec, (uint32_t)ntiles, bsum, info, mb, ms, flags);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(c.host_small + 4, sout, sizeof(ScanOut), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(c.host_small + 2, flags, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (reinterpret_cast<uint32_t*>(c.host_small + 2)[0]) return ALPACC_REJECT;
  const ScanOut so = *reinterpret_cast<const ScanOut*>(c.host_small + 4);
  const uint64_t np = (uint64_t)so.np;
  const uint64_t nodes = tree ? np - n : 0;
  if (so.nb > 0x7FFFFFFF || so.np > 0x7FFFFFFF)
    return fail(ALPACC_ERR_LIMIT, "derivation beyond 2^31 entries");
  // list sizes: open pushes + unmatched pops, bounded by the stream lengths
  if ((r = c.ps_b.reserve((size_t)so.nb * 4 + 64)) || (r = c.ps_c.reserve((size_t)(so.np + 1) * 16 + 64)) ||
      (r = c.tr_prod.reserve(nodes + 16)) || (r = c.tr_parent.reserve(nodes * 8 + 16)) ||
      (r = c.tr_tokpar.reserve(n * 8 + 16)) || (!tree && (r = c.ps_d.reserve(np + 16))))
    return r;
  uint16_t* ob_list = c.ps_b.as<uint16_t>();
  uint16_t* ub_list = ob_list + so.nb + 8;
  int64_t* os_list = c.ps_c.as<int64_t>();
  int64_t* us_list = os_list + so.np + 1;
  TileTree tb{}, ts{};
  if ((r = build_tile_tree(mb, ntiles, reinterpret_cast<int32_t*>(base + off_tb), st, &tb)) ||
      (r = build_tile_tree(ms, ntiles, reinterpret_cast<int32_t*>(base + off_ts), st, &ts)))
    return r;
  Resolve rs{pid, m, info, reinterpret_cast<const uint4*>(c.d_pair_rec), c.d_prod_info, g->nproductions, bcap, scap, ob_list, ub_list, os_list, us_list, c.tr_prod.as<uint8_t>(),
             c.tr_parent.as<int64_t>(), c.tr_tokpar.as<int64_t>(), tree ? nullptr : c.ps_d.as<uint8_t>(),
             tree ? 1 : 0, flags};
  if (!c.resolve_lds_set || c.resolve_lds_set < lds_b) {
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(tile_resolve),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_b));
    c.resolve_lds_set = lds_b;
  }
  hipLaunchKernelGGL(tile_resolve, dim3((uint32_t)ntiles), dim3(RWG), lds_b, st, rs);
  Far fr{info, tb, ts, ob_list, ub_list, os_list, us_list, c.tr_parent.as<int64_t>(), c.tr_tokpar.as<int64_t>(),
         tree ? 1 : 0, flags};
  hipLaunchKernelGGL(far_resolve, dim3((uint32_t)ntiles), dim3(TWG), 0, st, fr);
  HIP_TRY(hipGetLastError());
  if (c.timing) HIP_TRY(hipEventRecord(c.ev[3], st));
  HIP_TRY(hipMemcpyAsync(c.host_small + 2, flags, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (c.timing) {
    float ms_ = 0;
    HIP_TRY(hipEventElapsedTime(&ms_, c.ev[2], c.ev[3]));
    c.parse_ms += ms_;
    c.parse_launches++;
  }
  if (reinterpret_cast<uint32_t*>(c.host_small + 2)[0]) return ALPACC_REJECT;
  *num_prods = np;
  return ALPACC_VALID;
}

}  // namespace alpacc
