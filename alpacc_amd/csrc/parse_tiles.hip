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

constexpr int TWG = 256;         // lanes per workgroup
constexpr int TPL = 4;           // positions per lane
constexpr int TPOS = TWG * TPL;  // positions per tile
constexpr int MAXW = 5;          // brackets / productions per position handled in LDS

constexpr uint32_t BAD_PAIR = 1u, BAD_DEPTH = 2u, BAD_MATCH = 4u, BAD_COUNT = 8u, BAD_ROOT = 16u;
constexpr int64_t TOKEN_TARGET = 1ll << 62;

struct Pairs {
  const uint8_t* tok;
  uint64_t m;
  int q, k, base, empty_t, start_t, end_t;
  const uint16_t* window;
  const int32_t* spans;
  const int8_t* prof;  // 8 bytes per pair
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
        const int8_t* pf = a.prof + 8 * id;
        nb += a.spans[4 * id + 1];
        np += a.spans[4 * id + 3];
        nt += (uint8_t)pf[4];
        b = sm_cat(b, SM{pf[0], pf[1]});
        s = sm_cat(s, SM{pf[2], pf[3]});
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
  int64_t nb, np, nt;          // totals
  int64_t bend, send;          // final depths
  int32_t bmin, pad;           // global minimum bracket depth
};

__global__ __launch_bounds__(1024) void tile_scan(const TileRec* rec, uint32_t ntiles, TileInfo* info, int32_t* mb,
                                                   int32_t* ms, ScanOut* out, uint64_t ntokens, uint32_t* flags) {
  // per lane: a contiguous run of tiles; 9 running sums
  constexpr int NV = 9;
  __shared__ int64_t wsum[NV][16];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const uint32_t per = (ntiles + 1023) / 1024;
  const uint32_t lo = tid * per, hi = lo + per < ntiles ? lo + per : ntiles;
  int64_t v[NV] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t t = lo; t < hi; ++t) {
    TileRec r = rec[t];
    v[0] += r.nb;
    v[1] += r.np;
    v[2] += r.nt;
    v[3] += r.bs;
    v[4] += r.ss;
    v[5] += r.bs - r.bm;  // open pushes at the tile end
    v[6] += -r.bm;        // unmatched pops
    v[7] += r.ss - r.sm;
    v[8] += -r.sm;
  }
  int64_t inc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    inc[k] = (int64_t)wave_scan_add64((uint64_t)v[k]);
    if (lane == 63) wsum[k][wid] = inc[k];
  }
  __syncthreads();
  int64_t pre[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    int64_t o = 0;
    for (uint32_t w = 0; w < wid; ++w) o += wsum[k][w];
    pre[k] = o + inc[k] - v[k];
  }
  int32_t gmin = 0x7FFFFFFF;
  for (uint32_t t = lo; t < hi; ++t) {
    TileRec r = rec[t];
    TileInfo ti;
    ti.b_off = pre[0];
    ti.p_off = pre[1];
    ti.t_off = pre[2];
    ti.bstart = (int32_t)pre[3];
    ti.sstart = (int32_t)(pre[4] + 1);  // the virtual root slot
    ti.bmin = r.bm;
    ti.smin = r.sm;
    ti.ob_off = (int32_t)pre[5];
    ti.ub_off = (int32_t)pre[6];
    ti.os_off = (int32_t)pre[7];
    ti.us_off = (int32_t)pre[8];
    info[t] = ti;
    mb[t] = ti.bstart + r.bm;
    ms[t] = ti.sstart + r.sm;
    gmin = gmin < ti.bstart + r.bm ? gmin : ti.bstart + r.bm;
    pre[0] += r.nb;
    pre[1] += r.np;
    pre[2] += r.nt;
    pre[3] += r.bs;
    pre[4] += r.ss;
    pre[5] += r.bs - r.bm;
    pre[6] += -r.bm;
    pre[7] += r.ss - r.sm;
    pre[8] += -r.sm;
  }
  if (gmin < 0) atomicOr(flags, BAD_DEPTH);
  if (tid == 1023) {
    ScanOut o;
    o.nb = pre[0];
    o.np = pre[1];
    o.nt = pre[2];
    o.bend = pre[3];
    o.send = pre[4];
    o.bmin = 0;
    o.pad = 0;
    *out = o;
    if (pre[3] != 0) atomicOr(flags, BAD_DEPTH);
    if (ntokens != ~0ull && (uint64_t)pre[2] != ntokens) atomicOr(flags, BAD_COUNT);
  }
}

// ---------------------------------------------------------------- pass B

struct Resolve {
  const uint16_t* pid;
  uint64_t m;
  const TileInfo* info;
  const int32_t* spans;
  const uint16_t* stacks;
  const uint8_t* prods;
  const uint8_t* arity;
  const uint8_t* isterm;
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

constexpr int CAP = TPOS * MAXW;  // items per tile in LDS (per stream)
constexpr int CH = 32;            // chunk width of the LDS min index

// Last k < j with a[k] < v (STRICT) or a[k] <= v; -1 if none.  `cm` holds
// the minimum of every CH-wide chunk.
template <bool STRICT>
__device__ __forceinline__ int lds_prev(const int16_t* a, const int16_t* cm, int j, int v) {
  auto ok = [&](int x) { return STRICT ? x < v : x <= v; };
  int lo = j & ~(CH - 1);
  for (int k = j - 1; k >= lo; --k)
    if (ok(a[k])) return k;
  for (int c = (j / CH) - 1; c >= 0; --c) {
    if (ok(cm[c])) {
      for (int k = c * CH + CH - 1; k >= c * CH; --k)
        if (ok(a[k])) return k;
    }
  }
  return -1;
}

__global__ __launch_bounds__(TWG) void tile_resolve(Resolve r) {
  __shared__ uint16_t bsym[CAP];
  __shared__ int16_t bdep[CAP];
  __shared__ uint8_t bhit[CAP];
  __shared__ uint8_t sprod[CAP];
  __shared__ int16_t se[CAP];
  __shared__ uint16_t stinc[CAP];
  __shared__ __align__(4) uint8_t scons[CAP];
  __shared__ int16_t bcm[CAP / CH], scm[CAP / CH];
  __shared__ int32_t wsum[5][TWG / WAVE];

  const uint32_t tile = blockIdx.x;
  const TileInfo ti = r.info[tile];
  const uint64_t i0 = (uint64_t)tile * TPOS + (uint64_t)threadIdx.x * TPL;
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;

  uint32_t ids[TPL];
  int32_t nb = 0, np = 0, bd = 0, sd = 0, nt = 0;
#pragma unroll
  for (int k = 0; k < TPL; ++k) {
    uint64_t i = i0 + k;
    ids[k] = i < r.m ? r.pid[i] : 0xFFFFu;
    if (ids[k] != 0xFFFFu) {
      nb += r.spans[4 * ids[k] + 1];
      np += r.spans[4 * ids[k] + 3];
    }
  }
  // exclusive offsets of this lane's brackets / slots inside the tile
  uint32_t ib = wave_scan_add32((uint32_t)nb), ip = wave_scan_add32((uint32_t)np);
  if (lane == 63) {
    wsum[0][wid] = (int32_t)ib;
    wsum[1][wid] = (int32_t)ip;
  }
  __syncthreads();
  int32_t ob = (int32_t)ib - nb, op = (int32_t)ip - np;
  int32_t tb = 0, tp = 0;
  for (int w = 0; w < TWG / WAVE; ++w) {
    if ((uint32_t)w < wid) {
      ob += wsum[0][w];
      op += wsum[1][w];
    }
    tb += wsum[0][w];
    tp += wsum[1][w];
  }
  __syncthreads();
  // expand: brackets with relative depth-before; slots with the relative
  // depth after their pop and the inclusive terminal count
  {
    // lane depth prefixes need the per-lane sums first
    int32_t lbd = 0, lsd = 0, lnt = 0;
#pragma unroll
    for (int k = 0; k < TPL; ++k) {
      if (ids[k] == 0xFFFFu) continue;
      int32_t s0 = r.spans[4 * ids[k]], sl = r.spans[4 * ids[k] + 1];
      int32_t p0 = r.spans[4 * ids[k] + 2], pl = r.spans[4 * ids[k] + 3];
      for (int j = 0; j < sl; ++j) lbd += (r.stacks[s0 + j] & 0x8000u) ? 1 : -1;
      for (int j = 0; j < pl; ++j) {
        uint8_t pr = r.prods[p0 + j];
        lsd += (int32_t)r.arity[pr] - 1;
        lnt += r.isterm[pr];
      }
    }
    uint32_t xb = wave_scan_add32((uint32_t)lbd), xs = wave_scan_add32((uint32_t)lsd),
             xt = wave_scan_add32((uint32_t)lnt);
    if (lane == 63) {
      wsum[2][wid] = (int32_t)xb;
      wsum[3][wid] = (int32_t)xs;
      wsum[4][wid] = (int32_t)xt;
    }
    __syncthreads();
    bd = (int32_t)xb - lbd;
    sd = (int32_t)xs - lsd;
    nt = (int32_t)xt - lnt;
    for (uint32_t w = 0; w < wid; ++w) {
      bd += wsum[2][w];
      sd += wsum[3][w];
      nt += wsum[4][w];
    }
    int32_t bo = ob, po = op;
#pragma unroll
    for (int k = 0; k < TPL; ++k) {
      if (ids[k] == 0xFFFFu) continue;
      int32_t s0 = r.spans[4 * ids[k]], sl = r.spans[4 * ids[k] + 1];
      int32_t p0 = r.spans[4 * ids[k] + 2], pl = r.spans[4 * ids[k] + 3];
      for (int j = 0; j < sl; ++j) {
        uint16_t b = r.stacks[s0 + j];
        bsym[bo] = b;
        bdep[bo] = (int16_t)bd;
        bhit[bo] = 0;
        bd += (b & 0x8000u) ? 1 : -1;
        ++bo;
      }
      for (int j = 0; j < pl; ++j) {
        uint8_t pr = r.prods[p0 + j];
        sprod[po] = pr;
        se[po] = (int16_t)(sd - 1);
        nt += r.isterm[pr];
        stinc[po] = (uint16_t)nt;
        scons[po] = 0;
        sd += (int32_t)r.arity[pr] - 1;
        ++po;
      }
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c * CH < tb; c += TWG) {
    int lo = c * CH, hi = lo + CH < tb ? lo + CH : tb;
    int mn = 0x7FFF;
    for (int k = lo; k < hi; ++k) mn = bdep[k] < mn ? bdep[k] : mn;
    bcm[c] = (int16_t)mn;
  }
  for (int c = threadIdx.x; c * CH < tp; c += TWG) {
    int lo = c * CH, hi = lo + CH < tp ? lo + CH : tp;
    int mn = 0x7FFF;
    for (int k = lo; k < hi; ++k) mn = se[k] < mn ? se[k] : mn;
    scm[c] = (int16_t)mn;
  }
  __syncthreads();

  // brackets: pops match the last earlier bracket with a smaller depth-before
  bool bad = false;
  for (int j = threadIdx.x; j < tb; j += TWG) {
    uint16_t b = bsym[j];
    if (b & 0x8000u) continue;
    int d = bdep[j];
    int k = lds_prev<true>(bdep, bcm, j, d);
    if (k >= 0) {
      bad |= bsym[k] != (uint16_t)(b | 0x8000u);
      bhit[k] = 1;
    } else {
      r.ub_list[ti.ub_off + (-d)] = b;
    }
  }
  // slots: the parent is the owner of the last earlier slot with e <= e_j
  const int64_t p_off = ti.p_off, t_off = ti.t_off;
  if (r.deriv)
    for (int j = threadIdx.x; j < tp; j += TWG) r.deriv[p_off + j] = sprod[j];
  for (int j = threadIdx.x; r.tree && j < tp; j += TWG) {
    int v = se[j];
    int k = lds_prev<false>(se, scm, j, v);
    uint8_t pr = sprod[j];
    int64_t J = p_off + j, tinc = t_off + stinc[j];
    bool term = r.isterm[pr];
    int64_t target = term ? (TOKEN_TARGET | (tinc - 1)) : (J - tinc);
    if (!term) r.tree_prod[J - tinc] = pr;
    if (k >= 0) {
      int64_t K = p_off + k;
      int64_t pc = K - (t_off + stinc[k]);
      atomicOr(reinterpret_cast<uint32_t*>(scons + (k & ~3)), (1u << (v - se[k])) << (8 * (k & 3)));
      if (term)
        r.tok_par[tinc - 1] = pc;
      else
        r.tree_par[J - tinc] = pc;
    } else {
      r.us_list[ti.us_off + (-(v + 1))] = target;
    }
  }
  if (bad) atomicOr(r.flags, BAD_MATCH);
  __syncthreads();
  // pushes still open at the tile end, indexed by level above the tile minimum
  for (int j = threadIdx.x; j < tb; j += TWG) {
    uint16_t b = bsym[j];
    if ((b & 0x8000u) && !bhit[j]) r.ob_list[ti.ob_off + (bdep[j] + 1 - ti.bmin - 1)] = b;
  }
  for (int j = threadIdx.x; r.tree && j < tp; j += TWG) {
    uint8_t pr = sprod[j];
    int a = r.arity[pr];
    if (!a) continue;
    int64_t pc = p_off + j - (t_off + stinc[j]);
    uint32_t cons = scons[j];
    for (int l = 0; l < a; ++l)
      if (!((cons >> l) & 1u)) r.os_list[ti.os_off + (se[j] + 1 + l - ti.smin - 1)] = pc;
  }
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
  if (g->max_brackets > MAXW || g->max_prods > MAXW)
    return fail(ALPACC_ERR_LIMIT, "grammar has windows with more than %d brackets or productions", MAXW);
  const uint64_t m = n + 2;
  const uint64_t ntiles = (m + TPOS - 1) / TPOS;
  if (ntiles > 0x7FFFFFFFull) return fail(ALPACC_ERR_LIMIT, "too many tiles");
  int r;
  // workspace: pid | TileRec | TileInfo | mb | ms | trees | ScanOut
  const uint64_t tt = tree_scratch(ntiles);
  size_t off_pid = 0;
  size_t off_rec = (off_pid + m * 2 + 255) & ~size_t(255);
  size_t off_info = (off_rec + ntiles * sizeof(TileRec) + 255) & ~size_t(255);
  size_t off_mb = (off_info + ntiles * sizeof(TileInfo) + 255) & ~size_t(255);
  size_t off_ms = (off_mb + ntiles * 4 + 255) & ~size_t(255);
  size_t off_tb = (off_ms + ntiles * 4 + 255) & ~size_t(255);
  size_t off_ts = (off_tb + tt * 4 + 255) & ~size_t(255);
  size_t off_out = (off_ts + tt * 4 + 255) & ~size_t(255);
  size_t total = off_out + sizeof(ScanOut) + 256;
  if ((r = c.ps_a.reserve(total))) return r;
  uint8_t* base = c.ps_a.as<uint8_t>();
  uint16_t* pid = reinterpret_cast<uint16_t*>(base + off_pid);
  TileRec* rec = reinterpret_cast<TileRec*>(base + off_rec);
  TileInfo* info = reinterpret_cast<TileInfo*>(base + off_info);
  int32_t* mb = reinterpret_cast<int32_t*>(base + off_mb);
  int32_t* ms = reinterpret_cast<int32_t*>(base + off_ms);
  ScanOut* sout = reinterpret_cast<ScanOut*>(base + off_out);
  uint32_t* flags = reinterpret_cast<uint32_t*>(c.small.as<uint64_t>() + 2);
  if ((r = zero_words(flags, 4, st))) return r;

  if (c.timing) HIP_TRY(hipEventRecord(c.ev[2], st));
  Pairs a{d_tok, m, g->q, g->k, g->window_base, g->empty_t, g->start_t, g->end_t, c.d_window, c.d_spans,
          reinterpret_cast<const int8_t*>(c.d_profile)};
  hipLaunchKernelGGL(pair_profile, dim3((uint32_t)ntiles), dim3(TWG), 0, st, a, pid, rec, flags);
  hipLaunchKernelGGL(tile_scan, dim3(1), dim3(1024), 0, st, rec, (uint32_t)ntiles, info, mb, ms, sout,
                     tree ? n : ~0ull, flags);
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
  Resolve rs{pid, m, info, c.d_spans, c.d_stacks, c.d_productions, c.d_arity, c.d_isterm, ob_list, ub_list,
             os_list, us_list, c.tr_prod.as<uint8_t>(), c.tr_parent.as<int64_t>(), c.tr_tokpar.as<int64_t>(),
             tree ? nullptr : c.ps_d.as<uint8_t>(), tree ? 1 : 0, flags};
  hipLaunchKernelGGL(tile_resolve, dim3((uint32_t)ntiles), dim3(TWG), 0, st, rs);
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
