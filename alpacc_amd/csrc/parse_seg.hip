// LLP parser for gfx950 over a batch of token strings (a single string is a
// batch of one).
//
// Semantics are those of backends/c/parser.c: parse_test (:49-98) checks that
// every window has a pair and that the concatenated bracket strings are
// balanced and symbol-matched; compute_parents (:103-118) and compact_tree
// (:126-153) build the productions-only tree.  Both checks are matching
// problems over depth levels:
//   * brackets: a right bracket pops the level its depth-before names;
//   * derivation slots: every production pops one open child slot of its
//     parent and pushes `arity` new ones, so its parent is the owner of the
//     slot its pop takes (the first production pops a virtual root slot).
// The generator reduces every window to its net effect (pair records: pops
// reaching below the window, surviving pushes, and each production's parent
// when it lies inside the window), so a position is a handful of pops and
// pushes.  The match of a pop at level L is the last earlier position whose
// depth after its own pops is below L; it holds the push at that level.
//
// Every input owns whole tiles of PT framed positions ([start] tokens [end]);
// one lane per position.
//   A  seg_profile   pair ids and per-tile sums and (sum, min) depth
//                    profiles of both streams;
//   S  tile scans    tile offsets and starting depths (the depth scans run
//                    across inputs; each input's first tile adds the virtual
//                    root slot, and every check is relative to the input);
//   B  seg_resolve   per tile: block scans, a sparse table of the positions'
//                    post-pop depths in LDS, every pop finds its push inside
//                    the tile in O(log PT) LDS reads, or goes to the tile's
//                    unmatched list; pushes still open at the tile end go to
//                    its open list; outputs staged in LDS, written coalesced;
//   C  far_resolve   each unmatched pop finds the last earlier tile of its
//                    input whose minimum is below its level (a 32-ary min
//                    tree over tiles) and reads the open push there;
//   F  seg_finish    per input: balance, counts, offsets, verdict.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/alpacc_hip.h"
#include "device.hpp"
#include "lexer.hpp"
#include "parser.hpp"

namespace alpacc {

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return fail(ALPACC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

namespace {

// Tile shapes: PT framed positions per tile, APL positions per lane in
// pass B (so PT / APL lanes resolve a tile).  The default is (1024, 4);
// the context picks one of TILE_SHAPES (alpacc_hip_set_parse_shape and the
// parser autotuner).
constexpr int PT_DEFAULT = 1024, APL_DEFAULT = 4;
// Reject reasons.  parse_test (parser.c:49-98) rejects on pairs and brackets
// only; the tree checks add the terminal count (compact_tree, parser.c:
// 126-153) and the derivation-slot structure.
constexpr uint32_t BAD_PAIR = 1u, BAD_BDEPTH = 2u, BAD_MATCH = 4u, BAD_COUNT = 8u, BAD_ROOT = 16u, BAD_SDEPTH = 32u;
constexpr uint32_t PARSE_TEST_CHECKS = BAD_PAIR | BAD_BDEPTH | BAD_MATCH;
constexpr int64_t TOKEN_TARGET = 1ll << 62;
constexpr uint32_t NO_PAIR = 0xFFFFu;

// Batch and grammar, shared by the passes.
struct Seg {
  const uint8_t* tok;          // token ids of all inputs
  const uint8_t* tok_end;      // the end of the token buffer: reads stay below it
  const uint64_t* tok_off;     // [count + 1] first token of each input
  const uint32_t* tile_first;  // [count + 1] first tile of each input (inputs not parsed have none)
  const uint32_t* tile_seg;    // [ntiles] input of each tile
  // tiles of the launch: *nt_dev in device memory; ntiles is an upper bound
  // (equal to it unless the sequence is sized for a batch capacity) that
  // sizes the grids and the workspace
  const uint32_t* nt_dev;
  uint32_t count, ntiles;
  int q, k, base, empty_t, start_t, end_t;
  // per window code: the pair id (bits 0..15, NO_PAIR for a window without
  // one) and the head of its record: bracket pops and pushes, productions,
  // terminal productions, slot pops and pushes (a byte each, bits 16..63;
  // zero without a pair)
  const uint64_t* window;
  uint32_t window_bytes;  // its size (the bound of the buffer loads of pass A)
  // hashed window tables (the grammar's dense index would be too large):
  // the code of each of the 2^hbits slots, all ones when empty; null when
  // the window entries are a dense index
  const uint64_t* wkeys;
  int hbits;
  const uint4* rec;  // 64-byte pair records, 4 x uint4 each; record npairs is empty
  int npairs;
  uint32_t* flags;   // [count] BAD_* bits
};

// The tile's input and framed length.
struct TileOf {
  uint32_t seg, lt;  // input, tile index inside the input
  int64_t m;         // framed positions of the input: tokens + 2
  const uint8_t* tk;
};

__device__ __forceinline__ TileOf tile_of(const Seg& s, uint32_t tile) {
  TileOf t;
  t.seg = s.tile_seg[tile];
  t.lt = tile - s.tile_first[t.seg];
  const uint64_t a = s.tok_off[t.seg], b = s.tok_off[t.seg + 1];
  t.m = (int64_t)(b - a) + 2;
  t.tk = s.tok + a;
  return t;
}

// Window entry (pair id and record head) of framed position p, or NO_PAIR
// (32-bit positions: an input has fewer than 2^31 tokens).
__device__ __forceinline__ uint64_t pair_at(const Seg& s, const TileOf& t, int64_t p64) {
  const int32_t p = (int32_t)p64, m = (int32_t)t.m;
  uint32_t code = 0;
  const int w = s.q + s.k;
  for (int j = 0; j < w; ++j) {
    const int32_t f = p - s.q + j;
    uint32_t v;
    if (f <= 0 || f >= m - 1)
      v = f == 0 ? (uint32_t)s.start_t : (f == m - 1 ? (uint32_t)s.end_t : (uint32_t)s.empty_t);
    else
      v = t.tk[f - 1];
    if (v >= (uint32_t)s.base) return NO_PAIR;
    code = code * s.base + v;
  }
  return s.window[code];
}

// Window entry of framed position p in a hashed window table: the window
// code (64 bits), then linear probing from slot code * 0x9E3779B97F4A7C15
// (top hbits bits) to the code's slot (the table is at most half full, so
// an empty slot ends every search).
__device__ __forceinline__ uint64_t pair_hashed(const Seg& s, const TileOf& t, int64_t p64) {
  const int32_t p = (int32_t)p64, m = (int32_t)t.m;
  uint64_t code = 0;
  const int w = s.q + s.k;
  for (int j = 0; j < w; ++j) {
    const int32_t f = p - s.q + j;
    uint32_t v;
    if (f <= 0 || f >= m - 1)
      v = f == 0 ? (uint32_t)s.start_t : (f == m - 1 ? (uint32_t)s.end_t : (uint32_t)s.empty_t);
    else
      v = t.tk[f - 1];
    if (v >= (uint32_t)s.base) return NO_PAIR;
    code = code * (uint64_t)s.base + v;
  }
  const uint32_t mask = (1u << s.hbits) - 1u;
  uint32_t slot = (uint32_t)((code * 0x9E3779B97F4A7C15ull) >> (64 - s.hbits));
  for (;;) {
    const uint64_t k = s.wkeys[slot];
    if (k == code) return s.window[slot];
    if (k == ~0ull) return NO_PAIR;
    slot = (slot + 1) & mask;
  }
}

constexpr int MAXW = 8;  // widest window (q + k) the register path of pass A covers

// Window entries of framed positions p0 .. p0 + n - 1 (NO_PAIR past the
// input's end): the framed tokens all of them need are loaded first, one
// independent load each, then every window code is formed in registers.
// W bounds the window width (q + k <= W).
template <int W, int N>
__device__ __forceinline__ void window_ids(const Seg& s, const TileOf& t, int32_t p0, uint64_t* ids) {
  const int w = s.q + s.k;
  const int32_t m = (int32_t)t.m;
  uint32_t v[N + W - 1];
#pragma unroll
  for (int i = 0; i < N + W - 1; ++i) {
    const int32_t f = p0 - s.q + i;
    uint32_t x = (uint32_t)s.empty_t;
    if (f == 0)
      x = (uint32_t)s.start_t;
    else if (f == m - 1)
      x = (uint32_t)s.end_t;
    else if (i < N + w - 1 && f > 0 && f < m - 1)
      x = t.tk[f - 1];
    v[i] = x;
  }
#pragma unroll
  for (int r = 0; r < N; ++r) {
    uint32_t code = 0;
    bool ok = p0 + r < m;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      if (j < w) {
        ok &= v[r + j] < (uint32_t)s.base;
        code = code * (uint32_t)s.base + v[r + j];
      }
    }
    ids[r] = ok ? s.window[code] : (uint64_t)NO_PAIR;
  }
}

// (sum, min-prefix) monoid: a then b.
struct SM {
  int32_t s, m;
};
__device__ __forceinline__ SM sm_cat(SM a, SM b) { return {a.s + b.s, a.m < a.s + b.m ? a.m : a.s + b.m}; }

__device__ __forceinline__ SM wave_scan_sm(SM v) {
  // identity: sum 0, minimum above any prefix depth of a tile
  const U32x2 r = dpp_scan(U32x2{(uint32_t)v.s, (uint32_t)v.m}, U32x2{0u, 0x3FFFFFFFu}, Mov2{}, [](U32x2 a, U32x2 b) {
    const SM c = sm_cat(SM{(int32_t)a.a, (int32_t)a.b}, SM{(int32_t)b.a, (int32_t)b.b});
    return U32x2{(uint32_t)c.s, (uint32_t)c.m};
  });
  return SM{(int32_t)r.a, (int32_t)r.b};
}

struct TileRec {
  int32_t np, nt, bs, bm, ss, sm, pad0, pad1;
};

struct TileInfo {
  int64_t p_off, t_off;    // slots and terminal productions before the tile (all inputs)
  int64_t bstart, sstart;  // depths at the tile start (sstart includes its input's root slot)
  int32_t bmin, smin;      // tile-relative minima
  int32_t ob_off, ub_off, os_off, us_off;
};

// Per-input values written by the first tile of the input.
struct SegBase {
  int64_t p_off, t_off, bstart, sstart;
};

// Record fields: byte i of the first 8 bytes.
__device__ __forceinline__ uint32_t hb(const uint4& r0, int i) {
  return ((i < 4 ? r0.x : r0.y) >> (8 * (i & 3))) & 0xFFu;
}
// Fields of a 64-byte pair record held in registers as r0..r3 (the layout
// emit_hip.py writes): bracket symbol j (pops 0..3, pushes 4..7),
// production j, parent code j (0..7), owner e (0..23).
__device__ __forceinline__ uint32_t rec_br(const uint4& r0, const uint4& r1, int j) {
  const uint64_t w = j < 4 ? (((uint64_t)r0.w << 32) | r0.z) : (((uint64_t)r1.y << 32) | r1.x);
  return (uint32_t)(w >> (16 * (j & 3))) & 0xFFFFu;
}
__device__ __forceinline__ uint32_t rec_prod(const uint4& r1, int j) {
  return (uint32_t)((((uint64_t)r1.w << 32) | r1.z) >> (8 * j)) & 0xFFu;
}
__device__ __forceinline__ uint32_t rec_parent(const uint4& r2, int j) {
  return (uint32_t)((((uint64_t)r2.y << 32) | r2.x) >> (8 * j)) & 0xFFu;
}
__device__ __forceinline__ uint32_t rec_owner(const uint4& r2, const uint4& r3, int e) {
  const uint64_t w = e < 8 ? (((uint64_t)r2.w << 32) | r2.z)
                           : (e < 16 ? (((uint64_t)r3.y << 32) | r3.x) : (((uint64_t)r3.w << 32) | r3.z));
  return (uint32_t)(w >> (8 * (e & 7))) & 0xFFu;
}

// ---------------------------------------------------------------- pass A

// Pass A: one wave per tile, PPL consecutive positions per lane (16, or 8
// for tiles of 512), so every lane has PPL independent window-index loads
// and then PPL independent record-head loads in flight; a tile of more than
// 64 * PPL positions takes several such rounds.  The tile's sums and depth
// profiles are wave reductions (no workgroup barrier).
constexpr int AWG = 256;  // four tiles per workgroup
constexpr int AWAVES = AWG / WAVE;

// Pair ids of framed positions p0 .. p0 + N - 1 for windows of exactly W
// = q + k tokens.  Lanes whose windows lie inside the input's tokens (all
// but the lanes at an input's two ends) load their N + W - 1 tokens
// directly and roll the window code along; the others take window_ids.
template <int W, int N>
__device__ __forceinline__ void window_ids_exact(const Seg& s, const TileOf& t, int32_t p0, uint64_t* ids) {
  const int32_t f0 = p0 - s.q;  // framed index of the first token the windows read
  if (f0 > 0 && f0 + N + W - 2 < (int32_t)t.m - 1) {
    uint32_t v[N + W - 1];
    const uint8_t* src = t.tk + (f0 - 1);
    // the tokens' dwords: ND dword-aligned loads, realigned in registers,
    // when they lie inside the token buffer; else one load per token
    constexpr int ND = (N + W - 1 + 3) / 4 + 1;
    const uint32_t shift = (uint32_t)((uintptr_t)src & 3u);
    const uint8_t* lo = src - shift;
    if (lo >= s.tok && lo + 4 * ND <= s.tok_end) {
      typedef uint32_t u32a4 __attribute__((aligned(4)));
      const u32a4* d = reinterpret_cast<const u32a4*>(lo);
      uint32_t raw[ND];
#pragma unroll
      for (int k = 0; k < ND; ++k) raw[k] = d[k];
#pragma unroll
      for (int i = 0; i < N + W - 1; ++i) {
        const uint32_t wk = __builtin_amdgcn_alignbyte(raw[i / 4 + 1], raw[i / 4], shift);
        v[i] = (wk >> (8 * (i % 4))) & 0xFFu;
      }
    } else {
#pragma unroll
      for (int i = 0; i < N + W - 1; ++i) v[i] = src[i];
    }
    const uint32_t base = (uint32_t)s.base;
    uint32_t top = 1;  // base^(W - 1)
#pragma unroll
    for (int j = 1; j < W; ++j) top *= base;
    uint32_t code = 0, bad = 0;
#pragma unroll
    for (int j = 0; j < W - 1; ++j) {
      code = code * base + v[j];
      bad += v[j] >= base;
    }
    // the entries through a buffer resource: a 32-bit offset from the
    // table's base and no branch around the load; a window holding a token
    // that is no terminal (bad) has a code past the table or a wrong one,
    // and its load, bounds-checked by the hardware (zero past
    // window_bytes), is replaced by NO_PAIR
    const __amdgpu_buffer_rsrc_t wtab =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint64_t*>(s.window), 0, (int)s.window_bytes, 0x00020000);
#pragma unroll
    for (int r = 0; r < N; ++r) {
      code = code * base + v[r + W - 1];
      bad += v[r + W - 1] >= base;
      const auto e = __builtin_amdgcn_raw_buffer_load_b64(wtab, code * 8u, 0, 0);
      ids[r] = bad ? (uint64_t)NO_PAIR : ((uint64_t)e[1] << 32 | e[0]);
      code -= v[r] * top;
      bad -= v[r] >= base;
    }
  } else {
    window_ids<W, N>(s, t, p0, ids);
  }
}

// One round of pass A over PPL positions per lane from framed position p0:
// writes the pair ids, adds the lane's slot / terminal-production counts to
// np / nt, and returns the round's bracket and slot depth profiles over the
// whole wave (in every lane).
template <int W, int PPL>
__device__ __forceinline__ void profile_round(const Seg& s, const TileOf& t, int32_t p0, uint16_t* pid_at, int32_t& np,
                                              int32_t& nt, bool& bad, SM& wb, SM& ws) {
  // one load per position: the window entry holds the record's head
  uint64_t wr[PPL];
  if (W > 0) {
    window_ids_exact<(W > 0 ? W : 1), PPL>(s, t, p0, wr);
  } else if (s.wkeys) {
#pragma unroll
    for (int r = 0; r < PPL; ++r) wr[r] = p0 + r < t.m ? pair_hashed(s, t, p0 + r) : (uint64_t)NO_PAIR;
  } else if (s.q + s.k <= MAXW) {
    window_ids<MAXW, PPL>(s, t, p0, wr);
  } else {
#pragma unroll
    for (int r = 0; r < PPL; ++r) wr[r] = p0 + r < t.m ? pair_at(s, t, p0 + r) : (uint64_t)NO_PAIR;
  }
  uint32_t ids[PPL];
  SM b{0, 0}, sl{0, 0};
#pragma unroll
  for (int r = 0; r < PPL; ++r) {
    ids[r] = (uint32_t)wr[r] & 0xFFFFu;
    bad |= p0 + r < t.m && ids[r] == NO_PAIR;
    const uint32_t xx = (uint32_t)(wr[r] >> 16), xy = (uint32_t)(wr[r] >> 48);
    const int32_t bpop = xx & 0xFF, bpush = (xx >> 8) & 0xFF, spop = xy & 0xFF, spush = (xy >> 8) & 0xFF;
    np += (xx >> 16) & 0xFF;
    nt += xx >> 24;
    b = sm_cat(b, SM{bpush - bpop, -bpop});
    sl = sm_cat(sl, SM{spush - spop, -spop});
  }
  uint4* dst = reinterpret_cast<uint4*>(pid_at);
#pragma unroll
  for (int v = 0; v < PPL / 8; ++v)
    dst[v] = uint4{ids[8 * v] | ids[8 * v + 1] << 16, ids[8 * v + 2] | ids[8 * v + 3] << 16,
                   ids[8 * v + 4] | ids[8 * v + 5] << 16, ids[8 * v + 6] | ids[8 * v + 7] << 16};
  const SM ib = wave_scan_sm(b), is = wave_scan_sm(sl);
  wb = SM{__shfl(ib.s, WAVE - 1, WAVE), __shfl(ib.m, WAVE - 1, WAVE)};
  ws = SM{__shfl(is.s, WAVE - 1, WAVE), __shfl(is.m, WAVE - 1, WAVE)};
}

template <int W, int PT>
__device__ __forceinline__ void profile_tile(const Seg& s, uint32_t tile, uint32_t lane, TileRec* rec, uint16_t* pid) {
  constexpr int PPL = PT / WAVE < 16 ? PT / WAVE : 16, ROUNDS = PT / (WAVE * PPL);
  const TileOf t = tile_of(s, tile);
  int32_t np = 0, nt = 0;
  bool bad = false;
  SM tb, ts;  // the tile's profiles, round after round
  profile_round<W, PPL>(s, t, (int32_t)t.lt * PT + (int32_t)lane * PPL,
                        pid + (uint64_t)tile * PT + (uint64_t)lane * PPL, np, nt, bad, tb, ts);
  for (int round = 1; round < ROUNDS; ++round) {
    SM rb, rs;
    profile_round<W, PPL>(s, t, (int32_t)t.lt * PT + round * WAVE * PPL + (int32_t)lane * PPL,
                          pid + (uint64_t)tile * PT + (uint64_t)(round * WAVE + lane) * PPL, np, nt, bad, rb, rs);
    tb = sm_cat(tb, rb);
    ts = sm_cat(ts, rs);
  }
  const uint32_t cp = wave_scan_add32((uint32_t)np), ct = wave_scan_add32((uint32_t)nt);
  const uint64_t anybad = __ballot(bad);
  if (lane == WAVE - 1) {
    rec[tile] = TileRec{(int32_t)cp, (int32_t)ct, tb.s, tb.m, ts.s, ts.m, 0, 0};
    if (anybad) atomicOr(s.flags + t.seg, BAD_PAIR);
  }
}

// The thread index, opaque to the compiler: values derived from it are
// formed inside each tile's iteration instead of held in registers across
// the whole loop.
__device__ __forceinline__ uint32_t tid_per_tile() {
  uint32_t tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  return tid;
}

// One tile per wave: wave w of workgroup g takes tile AWAVES * g + w, or
// (STRIDE, for a grid sized before the tile count is known) the waves
// stride over the tiles.
// W = the grammar's window width q + k (1..6), or 0 for any width.
template <bool STRIDE, int W, int PT>
__global__ __launch_bounds__(AWG) void seg_profile(Seg s, TileRec* rec, uint16_t* pid) {
  const uint32_t nt = *s.nt_dev, lane = threadIdx.x & (WAVE - 1);
  const uint32_t w0 = blockIdx.x * AWAVES + threadIdx.x / WAVE;
  if (!STRIDE) {
    if (w0 < nt) profile_tile<W, PT>(s, w0, lane, rec, pid);
    return;
  }
  for (uint32_t tile = w0; tile < nt; tile += gridDim.x * AWAVES) profile_tile<W, PT>(s, tile, lane, rec, pid);
}

typedef void (*ProfileKernel)(Seg, TileRec*, uint16_t*);
template <bool STRIDE, int PT>
ProfileKernel profile_kernel(int w) {
  switch (w) {
    case 1: return seg_profile<STRIDE, 1, PT>;
    case 2: return seg_profile<STRIDE, 2, PT>;
    case 3: return seg_profile<STRIDE, 3, PT>;
    case 4: return seg_profile<STRIDE, 4, PT>;
    case 5: return seg_profile<STRIDE, 5, PT>;
    case 6: return seg_profile<STRIDE, 6, PT>;
    default: return seg_profile<STRIDE, 0, PT>;
  }
}
template <bool STRIDE>
ProfileKernel profile_kernel(int w, int pt) {
  return pt == 512 ? profile_kernel<STRIDE, 512>(w) : (pt == 2048 ? profile_kernel<STRIDE, 2048>(w)
                                                                  : profile_kernel<STRIDE, 1024>(w));
}

// ---------------------------------------------------------------- pass S

struct ScanOut {
  int64_t np, nt, bend, cend;  // totals: slots, terminal productions, bracket depth, slot depth
  int64_t ob, ub, os, us;      // list lengths
  // largest per-tile counts: nodes (slots less terminal productions),
  // terminal productions, slots; they size seg_resolve's LDS staging
  uint32_t tmax_node, tmax_term, tmax_slot;
  // set by seg_resolve when a tile holds more nodes or terminal productions
  // than the launch's staging (only a sequence sized for a capacity, with
  // staging learned from earlier batches, can see it): the launch's results
  // are then not valid and the batch runs again eagerly
  uint32_t overflow;
};

// The eight running sums over tiles: slots, terminal productions, bracket
// depth, slot depth (plus one root slot at each input's first tile), open
// pushes / unmatched pops of both streams.
constexpr int NV = 8;
constexpr int SB = 1024;  // tiles per scan block

__device__ __forceinline__ void tile_terms(const TileRec& r, bool first, int64_t* v) {
  v[0] = r.np;
  v[1] = r.nt;
  v[2] = r.bs;
  v[3] = r.ss + (first ? 1 : 0);
  v[4] = r.bs - r.bm;
  v[5] = -r.bm;
  v[6] = r.ss - r.sm;
  v[7] = -r.sm;
}

// Block-wide exclusive scan of NV int64 values (256 lanes); totals out.
__device__ __forceinline__ void block_scan_nv(int64_t* v, int64_t* tot) {
  __shared__ int64_t wsum[NV][4];
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const int64_t inc = (int64_t)wave_scan_add64((uint64_t)v[k]);
    if (lane == 63) wsum[k][wid] = inc;
    v[k] = inc - v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    int64_t o = 0, t = 0;
    for (uint32_t w = 0; w < 4; ++w) {
      if (w < wid) o += wsum[k][w];
      t += wsum[k][w];
    }
    v[k] += o;
    tot[k] = t;
  }
  __syncthreads();
}

__device__ __forceinline__ bool is_first(const Seg& s, uint32_t t) { return s.tile_first[s.tile_seg[t]] == t; }

__global__ __launch_bounds__(256) void tile_block_sums(Seg s, const TileRec* rec, int64_t* bsum, ScanOut* out) {
  int64_t v[NV] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t mx[3] = {0, 0, 0};  // per-tile maxima: nodes, terminal productions, slots
  const uint32_t t0 = blockIdx.x * SB + threadIdx.x * (SB / 256), nt = *s.nt_dev;
  for (int i = 0; i < SB / 256; ++i) {
    if (t0 + i >= nt) break;
    const TileRec r = rec[t0 + i];
    int64_t x[NV];
    tile_terms(r, is_first(s, t0 + i), x);
    for (int k = 0; k < NV; ++k) v[k] += x[k];
    mx[0] = max(mx[0], (uint32_t)(r.np - r.nt));
    mx[1] = max(mx[1], (uint32_t)r.nt);
    mx[2] = max(mx[2], (uint32_t)r.np);
  }
  __shared__ uint32_t wmx[3][256 / WAVE];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    for (int o = WAVE / 2; o > 0; o >>= 1) mx[k] = max(mx[k], (uint32_t)__shfl_xor((int)mx[k], o, WAVE));
    if ((threadIdx.x & (WAVE - 1)) == 0) wmx[k][threadIdx.x / WAVE] = mx[k];
  }
  int64_t tot[NV];
  block_scan_nv(v, tot);  // its barriers publish wmx
  if (threadIdx.x == 0) {
    for (int k = 0; k < NV; ++k) bsum[(uint64_t)blockIdx.x * NV + k] = tot[k];
    for (int k = 0; k < 3; ++k) {  // one atomic a block and count
      uint32_t m = 0;
      for (int w = 0; w < 256 / WAVE; ++w) m = max(m, wmx[k][w]);
      if (m) atomicMax(&out->tmax_node + k, m);
    }
  }
}

__global__ __launch_bounds__(256) void tile_block_scan(int64_t* bsum, uint32_t nblk, ScanOut* out) {
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
  if (threadIdx.x == 0) {  // the maxima are tile_block_sums' own
    out->np = carry[0];
    out->nt = carry[1];
    out->bend = carry[2];
    out->cend = carry[3];
    out->ob = carry[4];
    out->ub = carry[5];
    out->os = carry[6];
    out->us = carry[7];
  }
}

__global__ __launch_bounds__(256) void tile_downsweep(Seg s, const TileRec* rec, const int64_t* bsum, TileInfo* info,
                                                      SegBase* sb, int32_t* mb, int32_t* ms) {
  int64_t v[NV] = {0, 0, 0, 0, 0, 0, 0, 0};
  const uint32_t t0 = blockIdx.x * SB + threadIdx.x * (SB / 256), nt = *s.nt_dev;
  for (int i = 0; i < SB / 256; ++i) {
    if (t0 + i >= nt) break;
    int64_t x[NV];
    tile_terms(rec[t0 + i], is_first(s, t0 + i), x);
    for (int k = 0; k < NV; ++k) v[k] += x[k];
  }
  int64_t tot[NV];
  block_scan_nv(v, tot);
  for (int k = 0; k < NV; ++k) v[k] += bsum[(uint64_t)blockIdx.x * NV + k];
  for (int i = 0; i < SB / 256; ++i) {
    const uint32_t t = t0 + i;
    if (t >= nt) break;
    const TileRec r = rec[t];
    const bool first = is_first(s, t);
    TileInfo ti;
    ti.p_off = v[0];
    ti.t_off = v[1];
    ti.bstart = v[2];
    ti.sstart = v[3] + (first ? 1 : 0);
    ti.bmin = r.bm;
    ti.smin = r.sm;
    ti.ob_off = (int32_t)v[4];
    ti.ub_off = (int32_t)v[5];
    ti.os_off = (int32_t)v[6];
    ti.us_off = (int32_t)v[7];
    info[t] = ti;
    mb[t] = (int32_t)(ti.bstart + r.bm);
    ms[t] = (int32_t)(ti.sstart + r.sm);
    if (first) sb[s.tile_seg[t]] = SegBase{ti.p_off, ti.t_off, ti.bstart, ti.sstart};
    int64_t x[NV];
    tile_terms(r, first, x);
    for (int k = 0; k < NV; ++k) v[k] += x[k];
  }
}

// ---------------------------------------------------------------- pass B

constexpr int WLV = 6;            // in-wave sparse-table levels: blocks of 1..32 lanes
constexpr uint8_t TERM_SLOT = 0x80u;

struct Resolve {
  Seg s;
  const uint16_t* pid;  // [ntiles * positions per tile] pair ids from pass A
  const TileInfo* info;
  const SegBase* sb;
  int bcap, scap;      // per-lane stack room: positions per lane * (pops + pushes) of a window, brackets and slots
  int ncap, tcap;      // staged nodes / terminal productions per tile
  uint16_t* ob_list;   // open pushes (bracket symbols)
  uint16_t* ub_list;   // unmatched pops (bracket symbols)
  uint32_t* ub_tile;   // the tile of each unmatched pop (and of each unmatched slot below)
  uint32_t* us_tile;
  int64_t* os_list;    // open child slots (owner's compact index, all inputs)
  int64_t* us_list;    // unmatched slots (output target)
  uint8_t* tree_prod;
  int64_t* tree_par;
  int64_t* tok_par;
  uint8_t* deriv;  // flat derivation (parser-only mode) or null
  int tree;
  uint32_t* overflow;  // ScanOut::overflow of the launch
  // two-byte production ids (more than 255 productions): 8 u16 a pair
  // record, 16 bytes; productions are then staged and written as u16.
  // Null: the records' own bytes, written as u8.
  const uint4* prod16;
};

// Last lane l <= from whose minimum (tb[0][l]) is below L, or -1.  tb[k][l]
// is the minimum over lanes [max(w, l - 2^k + 1), l], w the first lane of
// l's wave, pw[l] the minimum over [w, l] and wmin[v] the minimum of wave v:
// the search descends inside from's wave when it holds such a lane, else
// inside the nearest earlier wave that does, from its last lane.
template <int RWG>
__device__ __forceinline__ int prev_below(const int16_t (*tb)[RWG], const int16_t* pw, const int32_t* wmin, int from,
                                          int32_t L) {
  auto descend = [&](int l) {  // a lane below L lies in [first lane of l's wave, l]
#pragma unroll
    for (int k = WLV - 1; k >= 0; --k)
      if (tb[k][l] >= L) l -= 1 << k;
    return l;
  };
  if (from < 0) return -1;
  if (tb[0][from] < L) return from;
  if (pw[from] < L) return descend(from);
  for (int v = from / WAVE - 1; v >= 0; --v)
    if (wmin[v] < L) return descend(v * WAVE + WAVE - 1);
  return -1;
}

// Block-wide exclusive scans of four int32 values.
template <int RNW>
__device__ __forceinline__ void block_scan4(int32_t* v, int32_t* tot, int32_t (*ws)[RNW]) {
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int32_t inc[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    inc[k] = (int32_t)wave_scan_add32((uint32_t)v[k]);
    if (lane == 63) ws[k][wid] = inc[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int32_t o = 0, t = 0;
    for (uint32_t w = 0; w < (uint32_t)RNW; ++w) {
      if (w < wid) o += ws[k][w];
      t += ws[k][w];
    }
    v[k] = o + inc[k] - v[k];
    tot[k] = t;
  }
}

// MBPOP..MSPUSH bound the window effects of the grammar (bracket pops and
// pushes, productions, surviving child slots); the host picks the smallest
// instantiation that covers them.  BT holds a bracket symbol: u8 when every
// symbol of the grammar is below 256 (half the stack's LDS), else u16.
//
// Every lane runs its APL positions' window effects in order on two short
// stacks in LDS -- bracket symbols, and open child slots as lane-relative
// owner indices -- interleaved by lane (entry e of lane t at e * RWG + t, so
// a wave's lanes touch distinct banks).  Pops that find the lane's stack
// empty are recorded from the top entry down.  A lane's open entries sit at
// consecutive levels just above its minimum, so a pop left over by a later
// lane at tile-relative level D reads entry D - min - 1 of the last earlier
// lane whose minimum is below D (a sparse table over the lanes' minima).
template <int MBPOP, int MBPUSH, int MNP, int MSPUSH, typename BT, int PT, int APL, bool P16>
__device__ __forceinline__ void resolve_tile(const Resolve& r, uint32_t tile, uint32_t tid) {
  constexpr int RWG = PT / APL;    // lanes per tile
  constexpr int RNW = RWG / WAVE;  // waves per tile
  constexpr bool NARROW = MBPUSH <= 4 && MSPUSH <= 4;
  // dynamic LDS: bracket stacks BT [bcap][RWG], slot stacks u8 [scap][RWG],
  // staged parents of nodes u16 [ncap] and of terminal productions u16
  // [tcap] (tile-local compact indices), staged productions u8 [ncap]
  extern __shared__ __align__(16) uint16_t dyn[];
  __shared__ int16_t tbb[WLV][RWG], tbs[WLV][RWG], pwb[RWG], pws[RWG];
  __shared__ uint32_t pkb[RWG], pks[RWG];  // (lane minimum << 16) | previous smaller lane
  __shared__ uint16_t l_cb[RWG], l_tb[RWG];  // each lane's first node / terminal production (tile-local)
  __shared__ int32_t ws[4][RNW];
  __shared__ int32_t wmin[2][RNW];
  const Seg& s = r.s;
  const uint32_t lane = tid & (WAVE - 1), wid = tid / WAVE;
  const int bcap = r.bcap, scap = r.scap;
  BT* const ball = reinterpret_cast<BT*>(dyn);
  BT* const bst = ball + tid;
  uint8_t* const sall = reinterpret_cast<uint8_t*>(ball + bcap * RWG);
  uint8_t* const sst = sall + tid;
  uint16_t* const s_par = reinterpret_cast<uint16_t*>(sall + (scap * RWG + 1) / 2 * 2);
  uint16_t* const s_tpar = s_par + r.ncap;
  uint8_t* const s_prod = reinterpret_cast<uint8_t*>(s_tpar + r.tcap);
  uint16_t* const s_prod16 = s_tpar + r.tcap;  // (two-byte ids: the same room, twice as long)
  const uint32_t seg = s.tile_seg[tile];
  const TileInfo ti = r.info[tile];
  const SegBase in = r.sb[seg];

  // phase 0: record heads of the lane's positions, lane sums and offsets
  uint32_t ids[APL];
  {
    const uint16_t* pp = r.pid + (uint64_t)tile * PT + (uint64_t)tid * APL;
    if (APL == 2) {
      const uint32_t pv = *reinterpret_cast<const uint32_t*>(pp);
      ids[0] = pv & 0xFFFFu;
      ids[1 % APL] = pv >> 16;
    } else {
#pragma unroll
      for (int v = 0; v < APL / 4; ++v) {
        const uint2 pv = reinterpret_cast<const uint2*>(pp)[v];
        ids[(4 * v) % APL] = pv.x & 0xFFFFu;
        ids[(4 * v + 1) % APL] = pv.x >> 16;
        ids[(4 * v + 2) % APL] = pv.y & 0xFFFFu;
        ids[(4 * v + 3) % APL] = pv.y >> 16;
      }
    }
  }
  // the records through a buffer resource: 32-bit offsets instead of a
  // 64-bit address per position
  const __amdgpu_buffer_rsrc_t rtab = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<uint4*>(s.rec), 0, (int)(((uint32_t)s.npairs + 1) * 64), 0x00020000);
  auto rec16 = [&](uint32_t off) -> uint4 {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(rtab, off, 0, 0);
    return uint4{v[0], v[1], v[2], v[3]};
  };
  uint32_t ro[APL];  // byte offset of each position's record
  uint4 h[APL];
  int32_t np = 0, nt = 0, bd = 0, sd = 0;
#pragma unroll
  for (int k = 0; k < APL; ++k) {
    ro[k] = (ids[k] == NO_PAIR ? (uint32_t)s.npairs : ids[k]) * 64u;
    h[k] = rec16(ro[k]);
    np += (h[k].x >> 16) & 0xFF;
    nt += h[k].x >> 24;
    bd += (int32_t)((h[k].x >> 8) & 0xFF) - (int32_t)(h[k].x & 0xFF);
    sd += (int32_t)((h[k].y >> 8) & 0xFF) - (int32_t)(h[k].y & 0xFF);
  }
  int32_t v[4] = {np, nt, bd, sd}, tot[4];
  block_scan4(v, tot, ws);
  const int32_t slot0 = v[0], term0 = v[1], bS = v[2], sS = v[3];
  const int32_t tile_np = tot[0], tile_nt = tot[1];
  // a tile beyond the staging (possible only with staging sized from
  // earlier batches) flags the launch and does nothing else: no stack, list
  // or output entry is written for it
  if ((r.tree ? tile_np - tile_nt : tile_np) > r.ncap || (r.tree ? tile_nt : 0) > r.tcap) {
    if (tid == 0) atomicOr(r.overflow, 1u);
    return;
  }
  const int32_t cb0 = slot0 - term0;  // tile-local compact index of the lane's first node
  l_cb[tid] = (uint16_t)cb0;
  l_tb[tid] = (uint16_t)term0;

  // phase 1: the lane's positions in order
  const int64_t c_off = ti.p_off - ti.t_off;  // compact index of the tile's first node (all inputs)
  const int64_t c_in = in.p_off - in.t_off;   // ... of the input's first node
  int bsp = 0, bun = 0, ssp = 0, sun = 0;
  int32_t bdep = 0, bmin = 0, sdep = 0, smin = 0;
  int32_t slot = slot0, tinc = term0;
  bool bad = false;
  // the rest of each record is held in registers while its position runs:
  // one 48-byte load at the position's start (its address waits for the
  // walk's state, so the compiler does not hoist the four loads together)
#pragma unroll
  for (int k = 0; k < APL; ++k) {
    uint32_t ra = ro[k];
    asm volatile("" : "+v"(ra) : "v"(slot));
    const uint4 r1 = rec16(ra + 16), r2 = rec16(ra + 32), r3 = rec16(ra + 48);
    const uint4 r0 = h[k];
    const uint32_t hx = r0.x, hy = r0.y;
    uint4 p16{0, 0, 0, 0};
    if (P16) p16 = r.prod16[ids[k] == NO_PAIR ? (uint32_t)s.npairs : ids[k]];
    const int nbpop = hx & 0xFF, nbpush = (hx >> 8) & 0xFF, npk = (hx >> 16) & 0xFF, ntk = hx >> 24;
    const int nspush = (hy >> 8) & 0xFF;
    const uint32_t tmask = (hy >> 16) & 0xFF;
    // the loops run to the grammar's maxima with constant field offsets;
    // iterations past the window's own counts do nothing
#pragma unroll
    for (int e = 0; e < MBPOP; ++e) {
      if (e < nbpop) {
        const BT sym = (BT)rec_br(r0, r1, e);
        if (bsp > 0) {
          bad |= bst[--bsp * RWG] != sym;
        } else {
          bst[(bcap - 1 - bun) * RWG] = sym;
          ++bun;
        }
      }
    }
    bdep -= nbpop;
    bmin = bdep < bmin ? bdep : bmin;
    // pushes of the narrow instantiations are written up to the maximum
    // with no test: an entry past the window's own pushes lies above the
    // stack top, where the next push writes before anything reads it (the
    // room then counts the instantiation's pushes for every position)
#pragma unroll
    for (int e = 0; e < MBPUSH; ++e)
      if (NARROW || e < nbpush) bst[(bsp + e) * RWG] = (BT)rec_br(r0, r1, 4 + e);
    bsp += nbpush;
    bdep += nbpush;
    const int32_t cbk = slot - tinc;  // tile-local compact index of the position's first node
    int32_t tseen = 0;
#pragma unroll
    for (int j = 0; j < MNP; ++j) {
      if (j < npk) {
        const uint32_t w16 = j < 2 ? p16.x : j < 4 ? p16.y : j < 6 ? p16.z : p16.w;
        const uint32_t pr = P16 ? (w16 >> (16 * (j & 1))) & 0xFFFFu : rec_prod(r1, j);
        const bool term = (tmask >> j) & 1u;
        const int32_t c = cbk + j - tseen;  // node (tile-local compact index)
        const int32_t tl = tinc + tseen;    // terminal production (tile-local)
        tseen += term;
        if (P16) {
          if (!r.tree)
            s_prod16[slot + j] = (uint16_t)pr;
          else if (!term)
            s_prod16[c] = (uint16_t)pr;
        } else if (!r.tree) {
          s_prod[slot + j] = (uint8_t)pr;
        } else if (!term) {
          s_prod[c] = (uint8_t)pr;
        }
        const uint32_t code = rec_parent(r2, j);
        int32_t par = cbk + (int32_t)(code & 0x7Fu);
        if (code & 0x80u) {
          if (ssp > 0) {
            par = cb0 + sst[--ssp * RWG];
          } else {
            sst[(scap - 1 - sun) * RWG] = term ? (uint8_t)(TERM_SLOT | (tl - term0)) : (uint8_t)(c - cb0);
            ++sun;
            par = -1;
          }
          --sdep;
          smin = sdep < smin ? sdep : smin;
        }
        if (r.tree && par >= 0) {
          if (term)
            s_tpar[tl] = (uint16_t)par;
          else
            s_par[c] = (uint16_t)par;
        }
      }
    }
#pragma unroll
    for (int e = 0; e < MSPUSH; ++e)
      if (NARROW || e < nspush) sst[(ssp + e) * RWG] = (uint8_t)(cbk - cb0 + rec_owner(r2, r3, e));
    ssp += nspush;
    sdep += nspush;
    slot += npk;
    tinc += ntk;
  }

  // phase 2: lane minima; in-wave sparse tables and prefix / suffix minima
  // by shuffles, wave minima in LDS; then every lane's previous strictly
  // smaller lane minimum (-1: none) -- lanes between a lane and that one
  // have minima >= its own
  const int32_t lbm = bS + bmin, lsm = sS + smin;
  int32_t pb = lbm, ps = lsm, sb = lbm, ss = lsm;
  tbb[0][tid] = (int16_t)lbm;
  tbs[0][tid] = (int16_t)lsm;
#pragma unroll
  for (int k = 1; k <= WLV; ++k) {  // the last step makes pb / ps the in-wave prefix minima
    const int hh = 1 << (k - 1);
    const int32_t a2 = __shfl_up(pb, hh, WAVE), c2 = __shfl_up(ps, hh, WAVE);
    if ((int)lane >= hh) {
      pb = a2 < pb ? a2 : pb;
      ps = c2 < ps ? c2 : ps;
    }
    if (k < WLV) {
      tbb[k][tid] = (int16_t)pb;
      tbs[k][tid] = (int16_t)ps;
    }
  }
  {  // in-wave inclusive suffix minima: a DPP prefix scan over the lanes in reverse
    const int rl = WAVE - 1 - (int)lane;
    const int32_t rb = wave_scan_min32(__shfl(sb, rl, WAVE)), rs = wave_scan_min32(__shfl(ss, rl, WAVE));
    sb = __shfl(rb, rl, WAVE);
    ss = __shfl(rs, rl, WAVE);
  }
  pwb[tid] = (int16_t)pb;
  pws[tid] = (int16_t)ps;
  if (lane == WAVE - 1) {
    wmin[0][wid] = pb;
    wmin[1][wid] = ps;
  }
  // the stacks, l_cb / l_tb, the tables and the wave minima are visible after this
  __syncthreads();
  int32_t bsuf = __shfl_down(sb, 1, WAVE), ssuf = __shfl_down(ss, 1, WAVE);
  if (lane == WAVE - 1) bsuf = ssuf = 0x7FFFFFFF;
  for (uint32_t v = wid + 1; v < (uint32_t)RNW; ++v) {
    bsuf = wmin[0][v] < bsuf ? wmin[0][v] : bsuf;
    ssuf = wmin[1][v] < ssuf ? wmin[1][v] : ssuf;
  }
  {
    const int pbl = prev_below(tbb, pwb, wmin[0], (int)tid - 1, lbm);
    const int psl = prev_below(tbs, pws, wmin[1], (int)tid - 1, lsm);
    pkb[tid] = (uint32_t)(uint16_t)(int16_t)lbm << 16 | (uint16_t)(int16_t)pbl;
    pks[tid] = (uint32_t)(uint16_t)(int16_t)lsm << 16 | (uint16_t)(int16_t)psl;
  }
  __syncthreads();

  // pops left over by this lane: a pop at tile-relative level D reads entry
  // D - min - 1 of the last earlier lane whose minimum is below D.  Depth
  // falls from pop to pop, so the lane only moves back, jumping from a lane
  // to its previous smaller one; the current lane's packed word stays in a
  // register.  Then the entries left open at the tile end.
  int l = (int)tid - 1;
  uint32_t w = l >= 0 ? pkb[l] : 0u;
#pragma clang loop unroll(disable) interleave(disable) vectorize(disable)
  for (int u = 0; u < bun; ++u) {
    const int32_t D = bS - u;
    const BT sym = bst[(bcap - 1 - u) * RWG];
    while (l >= 0 && (int32_t)(int16_t)(w >> 16) >= D) {
      l = (int16_t)(w & 0xFFFFu);
      if (l >= 0) w = pkb[l];
    }
    if (l >= 0)
      bad |= ball[(D - (int32_t)(int16_t)(w >> 16) - 1) * RWG + l] != sym;
    else
      r.ub_list[ti.ub_off + (-D)] = sym, r.ub_tile[ti.ub_off + (-D)] = tile;
  }
  // entry e sits at level lbm + 1 + e and stays open iff no later lane
  // falls below it: the first min(bsp, bsuf - lbm) entries
  const int64_t broom = (int64_t)bsuf - lbm;
  const int nob = broom < bsp ? (int)broom : bsp;
#pragma clang loop unroll(disable) interleave(disable) vectorize(disable)
  for (int e = 0; e < nob; ++e) r.ob_list[ti.ob_off + (lbm + e - ti.bmin)] = bst[e * RWG];
  if (r.tree) {
    l = (int)tid - 1;
    w = l >= 0 ? pks[l] : 0u;
#pragma clang loop unroll(disable) interleave(disable) vectorize(disable)
    for (int u = 0; u < sun; ++u) {
      const int32_t D = sS - u;
      const uint32_t tg = sst[(scap - 1 - u) * RWG];
      while (l >= 0 && (int32_t)(int16_t)(w >> 16) >= D) {
        l = (int16_t)(w & 0xFFFFu);
        if (l >= 0) w = pks[l];
      }
      const bool term = tg & TERM_SLOT;
      const int32_t idx = (term ? term0 : cb0) + (int32_t)(tg & 0x7Fu);
      if (l >= 0) {
        const uint16_t pc = (uint16_t)(l_cb[l] + sall[(D - (int32_t)(int16_t)(w >> 16) - 1) * RWG + l]);
        if (term)
          s_tpar[idx] = pc;
        else
          s_par[idx] = pc;
      } else {
        r.us_list[ti.us_off + (-D)] = term ? (TOKEN_TARGET | (ti.t_off + idx)) : (c_off + idx);
        r.us_tile[ti.us_off + (-D)] = tile;
      }
    }
    const int64_t sroom = (int64_t)ssuf - lsm;
    const int nos = sroom < ssp ? (int)sroom : ssp;
#pragma clang loop unroll(disable) interleave(disable) vectorize(disable)
    for (int e = 0; e < nos; ++e) r.os_list[ti.os_off + (lsm + e - ti.smin)] = c_off + cb0 + sst[e * RWG];
  }
  if (bad) atomicOr(s.flags + seg, BAD_MATCH);
  if (tid == 0 && ti.bstart + ti.bmin < in.bstart) atomicOr(s.flags + seg, BAD_BDEPTH);
  if (tid == 0 && ti.sstart + ti.smin < in.sstart - 1) atomicOr(s.flags + seg, BAD_SDEPTH);

  // coalesced write-out; entries whose parent lies before the tile are
  // written again by far_resolve, which runs after this kernel
  __syncthreads();
  if (r.tree) {
    const int32_t nn = tile_np - tile_nt;
#pragma clang loop unroll(disable) interleave(disable) vectorize(disable)
    for (int32_t i = tid; i < nn; i += RWG) {
      if (P16)
        reinterpret_cast<uint16_t*>(r.tree_prod)[c_off + i] = s_prod16[i];
      else
        r.tree_prod[c_off + i] = s_prod[i];
      r.tree_par[c_off + i] = c_off + s_par[i] - c_in;
    }
#pragma clang loop unroll(disable) interleave(disable) vectorize(disable)
    for (int32_t i = tid; i < tile_nt; i += RWG) r.tok_par[ti.t_off + i] = c_off + s_tpar[i] - c_in;
  } else {
#pragma clang loop unroll(disable) interleave(disable) vectorize(disable)
    for (int32_t i = tid; i < tile_np; i += RWG)
      if (P16)
        reinterpret_cast<uint16_t*>(r.deriv)[ti.p_off + i] = s_prod16[i];
      else
        r.deriv[ti.p_off + i] = s_prod[i];
  }
}

// One tile per workgroup, as seg_profile.
// P16: two-byte production ids (their own kernels, so that the one-byte
// grammars keep the registers of the one-byte path).
template <int MBPOP, int MBPUSH, int MNP, int MSPUSH, bool STRIDE, typename BT, int PT, int APL, bool P16>
__global__ __launch_bounds__(PT / APL) void seg_resolve(Resolve r) {
  const uint32_t nt = *r.s.nt_dev;
  if (!STRIDE) {
    if (blockIdx.x < nt) resolve_tile<MBPOP, MBPUSH, MNP, MSPUSH, BT, PT, APL, P16>(r, blockIdx.x, threadIdx.x);
    return;
  }
  for (uint32_t tile = blockIdx.x; tile < nt; tile += gridDim.x) {
    resolve_tile<MBPOP, MBPUSH, MNP, MSPUSH, BT, PT, APL, P16>(r, tile, tid_per_tile());
    __syncthreads();
  }
}

// The single-tile and striding kernels of one instantiation and tile shape,
// bracket stacks u8 (br8) or u16.
template <int MBPOP, int MBPUSH, int MNP, int MSPUSH, int PT, int APL, bool P16>
void resolve_kernels_shape(bool br8, void (**kern)(Resolve)) {
  kern[0] = br8 ? seg_resolve<MBPOP, MBPUSH, MNP, MSPUSH, false, uint8_t, PT, APL, P16>
                : seg_resolve<MBPOP, MBPUSH, MNP, MSPUSH, false, uint16_t, PT, APL, P16>;
  kern[1] = br8 ? seg_resolve<MBPOP, MBPUSH, MNP, MSPUSH, true, uint8_t, PT, APL, P16>
                : seg_resolve<MBPOP, MBPUSH, MNP, MSPUSH, true, uint16_t, PT, APL, P16>;
}
template <int MBPOP, int MBPUSH, int MNP, int MSPUSH, bool P16 = false>
void resolve_kernels(bool br8, int pt, int apl, void (**kern)(Resolve)) {
  if (pt == 512) resolve_kernels_shape<MBPOP, MBPUSH, MNP, MSPUSH, 512, 4, P16>(br8, kern);
  else if (pt == 2048) resolve_kernels_shape<MBPOP, MBPUSH, MNP, MSPUSH, 2048, 4, P16>(br8, kern);
  else if (apl == 2) resolve_kernels_shape<MBPOP, MBPUSH, MNP, MSPUSH, 1024, 2, P16>(br8, kern);
  else if (apl == 8) resolve_kernels_shape<MBPOP, MBPUSH, MNP, MSPUSH, 1024, 8, P16>(br8, kern);
  else resolve_kernels_shape<MBPOP, MBPUSH, MNP, MSPUSH, 1024, 4, P16>(br8, kern);
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
    const uint64_t lo = idx & ~31ull;
    for (uint64_t j = idx; j > lo;) {
      --j;
      if (a[j] < v) {
        uint64_t k = j;
        for (int D = L; D > 0; --D) {
          const int32_t* b = t.lv[D - 1];
          const uint64_t c0 = k * 32, c1 = c0 + 32 < t.n[D - 1] ? c0 + 32 : t.n[D - 1];
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
  Seg s;
  const uint32_t* overflow;  // set: seg_resolve skipped tiles, the lists are not this launch's
  const ScanOut* so;         // list lengths
  const TileInfo* info;
  const SegBase* sb;
  TileTree tb, ts;
  const uint16_t* ob_list;
  const uint16_t* ub_list;
  const int64_t* os_list;
  const int64_t* us_list;
  const uint32_t* ub_tile;
  const uint32_t* us_tile;
  int64_t* tree_par;
  int64_t* tok_par;
  int tree;
};

// One lane per pop left over at tile level, bracket pops then slot pops
// (the lists hold each pop's tile): the pop at depth D meets the open push
// of the last earlier tile of its input whose minimum is below D (32-ary
// min tree), which must carry its bracket symbol, or whose open child slot
// owns the production (the root slot, parent 0, when there is none).  A
// tile has few pops on JSON and thousands under deep nesting; every pop is
// its own lane either way.
__global__ __launch_bounds__(256) void far_resolve(Far f) {
  if (*f.overflow) return;  // the lists were not written: the batch runs again
  const int64_t nub = f.so->ub, nus = f.tree ? f.so->us : 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nub + nus; i += (int64_t)gridDim.x * 256) {
    if (i < nub) {
      const uint32_t tile = f.ub_tile[i];
      const TileInfo& ti = f.info[tile];
      const int32_t D = (int32_t)(ti.bstart - (i - ti.ub_off));  // depth before the pop
      const uint32_t seg = f.s.tile_seg[tile];
      const int64_t t2 = tile_prev_less(f.tb, tile, D);
      bool bad = t2 < (int64_t)f.s.tile_first[seg];
      if (!bad) {
        const TileInfo& o = f.info[t2];
        bad = f.ob_list[o.ob_off + (D - (int32_t)o.bstart - o.bmin - 1)] != f.ub_list[i];
      }
      if (bad) atomicOr(f.s.flags + seg, BAD_MATCH);
    } else {
      const int64_t j = i - nub;
      const uint32_t tile = f.us_tile[j];
      const TileInfo& ti = f.info[tile];
      const int32_t D = (int32_t)(ti.sstart - (j - ti.us_off));  // open child slots before the pop
      const uint32_t seg = f.s.tile_seg[tile];
      const int64_t target = f.us_list[j];
      const int64_t t2 = tile_prev_less(f.ts, tile, D);
      const SegBase& in = f.sb[seg];
      int64_t pc;
      if (t2 < (int64_t)f.s.tile_first[seg]) {
        if (D != (int32_t)in.sstart) atomicOr(f.s.flags + seg, BAD_ROOT);
        pc = 0;  // the root is its own parent
      } else {
        const TileInfo& o = f.info[t2];
        pc = f.os_list[o.os_off + (D - (int32_t)o.sstart - o.smin - 1)] - (in.p_off - in.t_off);
      }
      if (target & TOKEN_TARGET)
        f.tok_par[target & ~TOKEN_TARGET] = pc;
      else
        f.tree_par[target] = pc;
    }
  }
}

// Levels 1 and 2 of the two tile trees (bracket and slot depths) in one
// launch, blockIdx.y the tree: a workgroup takes 256 level-1 entries (8192
// tiles) and, when the trees have a level 2, the 8 level-2 entries over
// them (from LDS).
struct TreeLevels {
  const int32_t* in[2];
  int32_t* l1[2];
  int32_t* l2[2];  // null: no level 2
  uint64_t n0, n1, n2;
};
__global__ __launch_bounds__(256) void tiles_min_12(TreeLevels t) {
  __shared__ int32_t s_m[256];
  const uint32_t y = blockIdx.y;
  const uint64_t c = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  int32_t m = INT32_MAX;  // past the last entry: no effect on a minimum
  if (c < t.n1) {
    const int32_t* in = t.in[y];
    const uint64_t lo = c * 32, hi = lo + 32 < t.n0 ? lo + 32 : t.n0;
    m = in[lo];
    for (uint64_t i = lo + 1; i < hi; ++i) m = in[i] < m ? in[i] : m;
    t.l1[y][c] = m;
  }
  if (!t.l2[y]) return;  // the same in the whole workgroup
  s_m[threadIdx.x] = m;
  __syncthreads();
  const uint64_t c2 = (uint64_t)blockIdx.x * 8 + threadIdx.x;
  if (threadIdx.x < 8 && c2 < t.n2) {
    int32_t v = s_m[32 * threadIdx.x];
    for (uint32_t i = 1; i < 32; ++i) v = s_m[32 * threadIdx.x + i] < v ? s_m[32 * threadIdx.x + i] : v;
    t.l2[y][c2] = v;
  }
}

// One further level of both trees (blockIdx.y the tree).
__global__ __launch_bounds__(256) void tiles_min_level(const int32_t* in0, const int32_t* in1, uint64_t n_in,
                                                       int32_t* out0, int32_t* out1, uint64_t n_out) {
  const uint64_t c = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n_out) return;
  const int32_t* in = blockIdx.y ? in1 : in0;
  const uint64_t lo = c * 32, hi = lo + 32 < n_in ? lo + 32 : n_in;
  int32_t m = in[lo];
  for (uint64_t i = lo + 1; i < hi; ++i) m = in[i] < m ? in[i] : m;
  (blockIdx.y ? out1 : out0)[c] = m;
}

// The min trees over the n tile minima a (bracket depths) and b (slot
// depths), their upper levels in scratch_a / scratch_b: level L + 1 holds
// the minimum of every 32 entries of level L, up to a level of at most 32.
int build_tile_trees(const int32_t* a, const int32_t* b, uint64_t n, int32_t* scratch_a, int32_t* scratch_b,
                     hipStream_t st, TileTree* ta, TileTree* tb) {
  ta->levels = tb->levels = 1;
  ta->lv[0] = a;
  tb->lv[0] = b;
  ta->n[0] = tb->n[0] = n;
  int32_t *pa = scratch_a, *pb = scratch_b;
  for (uint64_t m = n; m > 32;) {
    if (ta->levels == MT_LEVELS) return fail(ALPACC_ERR_LIMIT, "tile tree too deep");
    const uint64_t nn = (m + 31) / 32;
    ta->lv[ta->levels] = pa;
    tb->lv[tb->levels] = pb;
    ta->n[ta->levels] = tb->n[tb->levels] = nn;
    ta->levels++;
    tb->levels++;
    pa += nn;
    pb += nn;
    m = nn;
  }
  const int levels = ta->levels;
  if (levels >= 2) {
    const bool two = levels >= 3;
    TreeLevels t{{a, b},
                 {const_cast<int32_t*>(ta->lv[1]), const_cast<int32_t*>(tb->lv[1])},
                 {two ? const_cast<int32_t*>(ta->lv[2]) : nullptr, two ? const_cast<int32_t*>(tb->lv[2]) : nullptr},
                 n, ta->n[1], two ? ta->n[2] : 0};
    hipLaunchKernelGGL(tiles_min_12, dim3((uint32_t)((ta->n[1] + 255) / 256), 2), dim3(256), 0, st, t);
  }
  for (int L = 3; L < levels; ++L)
    hipLaunchKernelGGL(tiles_min_level, dim3((uint32_t)((ta->n[L] + 255) / 256), 2), dim3(256), 0, st, ta->lv[L - 1],
                       tb->lv[L - 1], ta->n[L - 1], const_cast<int32_t*>(ta->lv[L]), const_cast<int32_t*>(tb->lv[L]),
                       ta->n[L]);
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

// ---------------------------------------------------------------- pass F

__global__ void seg_finish(Seg s, const TileInfo* info, const SegBase* sb, const ScanOut* so, int tree,
                           ParseResult* res) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.count) return;
  const uint32_t f0 = s.tile_first[i], f1 = s.tile_first[i + 1];
  ParseResult o{};
  if (f0 == f1) {  // not parsed
    o.flags = BAD_PAIR;
    res[i] = o;
    return;
  }
  const SegBase in = sb[i];
  int64_t p_end, t_end, b_end, c_end;
  if (f1 < *s.nt_dev) {
    const TileInfo& e = info[f1];
    p_end = e.p_off;
    t_end = e.t_off;
    b_end = e.bstart;
    c_end = e.sstart - 1;  // the next tile starts an input: drop its root slot
  } else {
    p_end = so->np;
    t_end = so->nt;
    b_end = so->bend;
    c_end = so->cend;
  }
  uint32_t fl = s.flags[i];
  if (b_end != in.bstart) fl |= BAD_BDEPTH;
  if (c_end != in.sstart - 1) fl |= BAD_SDEPTH;
  const int64_t ntok = (int64_t)(s.tok_off[i + 1] - s.tok_off[i]);
  if (t_end - in.t_off != ntok) fl |= BAD_COUNT;
  o.flags = tree ? fl : (fl & PARSE_TEST_CHECKS);
  o.slot_off = (uint64_t)in.p_off;
  o.slots = (uint64_t)(p_end - in.p_off);
  o.node_off = (uint64_t)(in.p_off - in.t_off);
  o.nodes = (uint64_t)((p_end - in.p_off) - (t_end - in.t_off));
  o.term_off = (uint64_t)in.t_off;
  res[i] = o;
}

// tile_seg[t] = the last input i with tile_first[i] <= t, for the launch's
// tiles.
__global__ void pseg_map(const uint32_t* tile_first, uint32_t count, const uint32_t* nt_dev, uint32_t* tile_seg) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= *nt_dev) return;
  uint32_t lo = 0, hi = count;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (tile_first[mid] <= t)
      lo = mid;
    else
      hi = mid;
  }
  tile_seg[t] = lo;
}

// The tiles of a batch whose token counts are known on the device only:
// input i gets ceil((tokens + 2) / PT) tiles unless skip[i] is set or i is
// at or past *live (padding); tile_first is their exclusive scan and
// *nt_dev the total.  One workgroup of PLAN_WG lanes, in rounds of
// PLAN_WG * PLAN_IPT inputs.
constexpr int PLAN_WG = 1024, PLAN_IPT = 4;
// A batch with more tokens than the sequence is sized for (tokens_cap, a
// capacity learned from earlier batches; 0: none beyond the bytes) or more
// tiles than ntiles_cap sets *overflow and leaves no tile: the batch runs
// again eagerly.  (The token bound is the one that matters: positions size
// the lists and the tree outputs.)
__global__ __launch_bounds__(PLAN_WG) void parse_plan(const uint64_t* tok_off, const uint32_t* skip,
                                                      const uint32_t* live, uint32_t count, uint32_t PT,
                                                      uint32_t* tile_first, uint32_t* nt_dev, uint32_t ntiles_cap,
                                                      uint64_t tokens_cap, uint32_t* overflow) {
  __shared__ uint32_t s_w[PLAN_WG / WAVE];
  const uint32_t tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE, n_live = *live;
  const uint32_t sh = __builtin_ctz(PT);  // a power of two
  uint32_t carry = 0;
  for (uint32_t r0 = 0; r0 < count; r0 += PLAN_WG * PLAN_IPT) {
    // framed positions ([|-] tokens [-|]) in whole tiles (every load
    // unconditional, from an index clamped to the last input, so that none
    // waits on another); an input's tokens fit 32 bits (inputs are below
    // 2^31 bytes)
    uint32_t t[PLAN_IPT], sum = 0;
#pragma unroll
    for (int k = 0; k < PLAN_IPT; ++k) {
      const uint32_t i = r0 + tid * PLAN_IPT + k, ic = i < count ? i : count - 1;
      const uint64_t a = tok_off[ic], b = tok_off[ic + 1];
      const uint32_t sk = skip[ic];
      t[k] = i < count && i < n_live && !sk ? ((uint32_t)(b - a) + 2 + PT - 1) >> sh : 0u;
      sum += t[k];
    }
    const uint32_t inc = wave_scan_add32(sum);
    if (lane == WAVE - 1) s_w[wid] = inc;
    __syncthreads();
    uint32_t pre = carry, tot = 0;
    for (uint32_t w = 0; w < PLAN_WG / WAVE; ++w) {
      if (w < wid) pre += s_w[w];
      tot += s_w[w];
    }
    pre += inc - sum;
#pragma unroll
    for (int k = 0; k < PLAN_IPT; ++k) {
      const uint32_t i = r0 + tid * PLAN_IPT + k;
      if (i < count) tile_first[i] = pre;
      pre += t[k];
    }
    carry += tot;
    __syncthreads();
  }
  // (carry is the same in every lane)
  const bool fits = carry <= ntiles_cap && (tokens_cap == 0 || tok_off[count] - tok_off[0] <= tokens_cap);
  if (!fits) {  // no input parsed
    __syncthreads();
    for (uint32_t i = tid; i < count; i += PLAN_WG) tile_first[i] = 0;
  }
  if (tid == 0) {
    tile_first[count] = fits ? carry : 0u;
    *nt_dev = fits ? carry : 0u;
    if (!fits) *overflow = 1u;
  }
}

}  // namespace

bool parse_scan_overflow(const Context& c) { return c.h_pres.as<ScanOut>()->overflow != 0; }

void parse_scan_maxima(const Context& c, bool tree, int* nodes, int* terms) {
  const ScanOut* so = c.h_pres.as<ScanOut>();
  *nodes = (int)(tree ? so->tmax_node : so->tmax_slot);
  *terms = tree ? (int)so->tmax_term : 0;
}

int parse_per_position(const GrammarDesc* g) {
  return std::max(std::max(1, std::max(g->max_bpush, g->max_spush)), std::max(g->max_bpop, g->max_spop));
}

bool parse_shape_ok(int pt, int apl) {
  return (pt == 1024 && (apl == 2 || apl == 4 || apl == 8)) || ((pt == 512 || pt == 2048) && apl == 4);
}

uint64_t parse_max_positions(const GrammarDesc* g) {
  return (0x7FFFFFFFull - 1) / (uint64_t)(parse_per_position(g) + 1);
}

uint32_t bracket_symbol_max(const GrammarDesc* g) {  // bytes 8..23 of each pair record, u16
  uint32_t mx = 0;
  for (int i = 0; i <= g->npairs; ++i)
    for (int j = 8; j < 24; j += 2)
      mx = std::max<uint32_t>(mx, g->pair_rec[(size_t)i * 64 + j] | (uint32_t)g->pair_rec[(size_t)i * 64 + j + 1] << 8);
  return mx;
}

int parse_batch_on_device(Context& c, const uint8_t* d_tok, const uint64_t* d_tok_off, const uint64_t* h_tok_off,
                          const uint32_t* h_skip, uint64_t count, bool tree, hipStream_t st, const BatchCaps* caps,
                          const uint32_t* d_skip, const uint32_t* d_live) {
  const GrammarDesc* g = c.g;
  if (!g->has_parser) return fail(ALPACC_ERR_ARG, "grammar '%s' has no production rules: lexer mode only", g->name);
  if (count == 0 || count > 0xFFFFFFFEull) return fail(ALPACC_ERR_ARG, "batch of %llu inputs", (unsigned long long)count);
  if (caps && (count != caps->count || !d_skip || !d_live))
    return fail(ALPACC_ERR_ARG, "batch beyond the capacity its launch sequence is sized for");
  // the context's tile shape: positions per tile and per lane of pass B
  const int PT = c.parse_pt, APL = c.parse_apl, RWG = PT / APL;
  if (!parse_shape_ok(PT, APL)) return fail(ALPACC_ERR_ARG, "parser tile shape %d / %d", PT, APL);
  const int ncap = PT * g->max_prods, tcap = PT * g->max_nt;
  // the resolve kernel's instantiation: window maxima it is compiled for
  // (two-byte production ids: the widest only)
  int ib[4];
  if (g->production_bytes == 2)
    ib[0] = 4, ib[1] = 4, ib[2] = 8, ib[3] = 24;
  else if (g->max_bpop <= 2 && g->max_bpush <= 3 && g->max_prods <= 4 && g->max_spush <= 3)
    ib[0] = 2, ib[1] = 3, ib[2] = 4, ib[3] = 3;
  else if (g->max_bpop <= 3 && g->max_bpush <= 4 && g->max_prods <= 4 && g->max_spush <= 4)
    ib[0] = 3, ib[1] = 4, ib[2] = 4, ib[3] = 4;
  else
    ib[0] = 4, ib[1] = 4, ib[2] = 8, ib[3] = 24;
  // lane stack room: a position's pops and pushes, the pushes counted to
  // the narrow instantiations' maxima (those kernels write that many)
  const bool narrow = ib[3] <= 4;
  const int bcap = APL * (g->max_bpop + (narrow ? ib[1] : g->max_bpush));
  const int scap = APL * (g->max_spop + (narrow ? ib[3] : g->max_spush));
  if (c.resolve_br8 < 0) c.resolve_br8 = bracket_symbol_max(g) < 256 ? 1 : 0;
  const bool br8 = c.resolve_bw ? c.resolve_bw == 1 : c.resolve_br8 == 1;
  // seg_resolve's dynamic LDS with node / terminal staging for n / t entries
  // (bracket stacks of bw bytes an entry)
  auto resolve_lds = [&](int n, int t, int bw) {
    return (size_t)bcap * RWG * bw + ((size_t)scap * RWG + 1) / 2 * 2 + (size_t)(n + t) * 2 +
           (size_t)n * g->production_bytes + 16;
  };
  const size_t lds_b = resolve_lds(ncap, tcap, br8 ? 1 : 2);
  // tile-local depths are int16, compact indices u16, lane-relative slot
  // owners 7 bits
  if ((int64_t)PT * 8 > 0x7FFF || ncap > 0xFFFF || APL * g->max_prods > 0x7F || lds_b > 96 * 1024)
    return fail(ALPACC_ERR_LIMIT, "grammar windows too large for the parser tiles");
  int r;
  // tiles per input: framed positions = tokens + 2; inputs skipped get none.
  // At capacity every input holds at most caps->bytes tokens in all.
  uint64_t ntiles = 0, positions = 0;
  uint32_t* h_tf = nullptr;
  if (caps) {
    const uint64_t tok = caps->tokens && caps->tokens < caps->bytes ? caps->tokens : caps->bytes;
    positions = tok + 2 * count;
    ntiles = (tok + count * (PT + 1)) / PT + 1;
    if ((r = c.h_pmeta.reserve(64))) return r;
  } else {
    for (uint64_t i = 0; i < count; ++i) {
      if (h_skip && h_skip[i]) continue;
      const uint64_t m = h_tok_off[i + 1] - h_tok_off[i] + 2;
      ntiles += (m + PT - 1) / PT;
      positions += m;
    }
    // host metadata: tile_first u32[count + 1], then the tile count
    if ((r = c.h_pmeta.reserve((count + 2) * 4 + 64))) return r;
    h_tf = c.h_pmeta.as<uint32_t>();
    uint32_t tt = 0;
    for (uint64_t i = 0; i < count; ++i) {
      h_tf[i] = tt;
      if (h_skip && h_skip[i]) continue;
      tt += (uint32_t)((h_tok_off[i + 1] - h_tok_off[i] + 2 + PT - 1) / PT);
    }
    h_tf[count] = tt;
    h_tf[count + 1] = tt;
  }
  // depths, list offsets and the tiles' depth minima move by at most one
  // window's pushes or pops per position: positions times the largest of
  // those bounds every 31-bit field of the scans, accepted input or not
  const uint64_t per_pos = (uint64_t)parse_per_position(g);
  if (ntiles > 0x7FFFFFFFull || positions * per_pos + count > 0x7FFFFFFFull)
    return fail(ALPACC_ERR_LIMIT, "batch of %llu framed positions exceeds the parser's 31-bit depths",
                (unsigned long long)positions);

  // device workspace: tile_first (+ tile count) | tile_seg | TileRec |
  // TileInfo | SegBase | mb | ms | trees | block sums | ScanOut |
  // ParseResult | flags | pair ids
  const uint64_t tts = tree_scratch(ntiles ? ntiles : 1);
  const uint32_t nsb = (uint32_t)((ntiles + SB - 1) / SB);
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t off_tf = 0;
  const size_t off_ts = al(off_tf + (count + 2) * 4);
  const size_t off_rec = al(off_ts + ntiles * 4);
  const size_t off_info = al(off_rec + ntiles * sizeof(TileRec));
  const size_t off_sb = al(off_info + ntiles * sizeof(TileInfo));
  const size_t off_mb = al(off_sb + count * sizeof(SegBase));
  const size_t off_ms = al(off_mb + ntiles * 4);
  const size_t off_tb = al(off_ms + ntiles * 4);
  const size_t off_tsx = al(off_tb + tts * 4);
  const size_t off_bs = al(off_tsx + tts * 4);
  const size_t off_out = al(off_bs + (size_t)(nsb ? nsb : 1) * NV * 8);
  // the scan's totals, then the per-input results: one copy brings both to
  // the host mirror (h_pres, the same layout)
  const size_t off_res = al(off_out + sizeof(ScanOut)), res_head = off_res - off_out;
  const size_t off_fl = al(off_res + count * sizeof(ParseResult));
  const size_t off_pid = al(off_fl + count * 4);
  const size_t total = off_pid + ntiles * PT * 2 + 64;
  if ((r = c.ps_a.reserve(total)) || (r = c.h_pres.reserve(res_head + count * sizeof(ParseResult) + 64)))
    return r;
  uint8_t* base = c.ps_a.as<uint8_t>();
  uint32_t* d_tf = reinterpret_cast<uint32_t*>(base + off_tf);
  Seg s{};
  s.tok = d_tok;
  // the token buffer holds the batch's tokens (at capacity: its bytes)
  s.tok_end = d_tok + (caps ? caps->bytes : h_tok_off[count]);
  s.tok_off = d_tok_off;
  s.tile_first = d_tf;
  s.tile_seg = reinterpret_cast<const uint32_t*>(base + off_ts);
  s.nt_dev = d_tf + count + 1;
  s.count = (uint32_t)count;
  s.ntiles = (uint32_t)ntiles;
  s.q = g->q;
  s.k = g->k;
  s.base = g->window_base;
  s.empty_t = g->empty_t;
  s.start_t = g->start_t;
  s.end_t = g->end_t;
  s.window = c.d_wrec;
  s.window_bytes = (uint32_t)std::min<uint64_t>((uint64_t)g->nwindow * 8, 0xFFFFFFF8ull);
  s.wkeys = g->window_hash ? c.d_wkeys : nullptr;
  s.hbits = g->window_hash;
  s.rec = reinterpret_cast<const uint4*>(c.d_pair_rec);
  s.npairs = g->npairs;
  s.flags = reinterpret_cast<uint32_t*>(base + off_fl);
  TileRec* rec = reinterpret_cast<TileRec*>(base + off_rec);
  TileInfo* info = reinterpret_cast<TileInfo*>(base + off_info);
  SegBase* sb = reinterpret_cast<SegBase*>(base + off_sb);
  int32_t* mb = reinterpret_cast<int32_t*>(base + off_mb);
  int32_t* ms = reinterpret_cast<int32_t*>(base + off_ms);
  int64_t* bsum = reinterpret_cast<int64_t*>(base + off_bs);
  ScanOut* sout = reinterpret_cast<ScanOut*>(base + off_out);
  ParseResult* res = reinterpret_cast<ParseResult*>(base + off_res);

  if ((r = zero_ranges({{s.flags, count * 4}, {sout, sizeof(ScanOut)}}, st))) return r;
  if (caps) {
    hipLaunchKernelGGL(parse_plan, dim3(1), dim3(PLAN_WG), 0, st, d_tok_off, d_skip, d_live, s.count, (uint32_t)PT, d_tf,
                       d_tf + count + 1, (uint32_t)ntiles,
                       caps->tokens && caps->tokens < caps->bytes ? caps->tokens : 0ull, &sout->overflow);
  } else {
    HIP_TRY(hipMemcpyAsync(d_tf, h_tf, (count + 2) * 4, hipMemcpyHostToDevice, st));
  }
  KernelTimer& tm = c.timer;
  const bool timed = c.timing && !caps;
  if (timed) HIP_TRY(hipEventRecord(c.ev[2], st));
  // grids: one workgroup per tile when the tile count is known, else as
  // many as the device holds at once, striding over the tiles
  const uint32_t g_map = (uint32_t)((ntiles + 255) / 256);
  // A sequence sized for a capacity launches a workgroup (pass A: a wave)
  // per tile the capacity can hold, and those past the actual count return
  // at once: cheaper than the striding instantiations, which hold more
  // registers across their loop (fewer waves per SIMD).  Only a capacity
  // beyond 2^20 tiles strides.
  const bool stride = caps && ntiles > (1u << 20);
  const uint32_t g_prof = (uint32_t)(stride ? 4096 : (ntiles + AWAVES - 1) / AWAVES);
  const uint32_t g_res = (uint32_t)(stride ? 2048 : ntiles);
  // far_resolve: one lane per leftover pop, at most 4096 workgroups striding (the
  // grid is set from the lists' sizes below)
  if (ntiles) {
    hipLaunchKernelGGL(pseg_map, dim3(g_map), dim3(256), 0, st, d_tf, s.count, s.nt_dev,
                       const_cast<uint32_t*>(s.tile_seg));
    if (!caps) HIP_TRY(tm.begin(st, "seg_profile"));
    // a hashed window table takes the kernel for any width (its lookups
    // probe the table)
    const int pw = g->window_hash ? 0 : g->q + g->k;
    hipLaunchKernelGGL(stride ? profile_kernel<true>(pw, PT) : profile_kernel<false>(pw, PT),
                       dim3(g_prof), dim3(AWG), 0, st, s, rec, reinterpret_cast<uint16_t*>(base + off_pid));
    if (!caps) HIP_TRY(tm.end(st));
    if (!caps) HIP_TRY(tm.begin(st, "tile_scan"));
    hipLaunchKernelGGL(tile_block_sums, dim3(nsb), dim3(256), 0, st, s, rec, bsum, sout);
    hipLaunchKernelGGL(tile_block_scan, dim3(1), dim3(256), 0, st, bsum, nsb, sout);
    hipLaunchKernelGGL(tile_downsweep, dim3(nsb), dim3(256), 0, st, s, rec, bsum, info, sb, mb, ms);
    if (!caps) HIP_TRY(tm.end(st));
  }
  HIP_TRY(hipGetLastError());
  // list and output sizes: the scan's totals, or at capacity the most that
  // the framed positions can produce
  ScanOut so;
  if (caps) {
    const int64_t pos = (int64_t)positions;
    so = ScanOut{pos * g->max_prods, pos * g->max_nt, 0, 0,
                 pos * g->max_bpush, pos * g->max_bpop, pos * g->max_spush, pos * g->max_spop};
  } else {
    HIP_TRY(hipMemcpyAsync(c.h_pres.ptr, sout, sizeof(ScanOut), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(tm.collect());
    so = *c.h_pres.as<ScanOut>();
  }
  const uint64_t np = (uint64_t)so.np, nterm = (uint64_t)so.nt;
  if (!caps && (so.np > 0x7FFFFFFF || so.ob > 0x7FFFFFFF || so.ub > 0x7FFFFFFF || so.os > 0x7FFFFFFF ||
                so.us > 0x7FFFFFFF))
    return fail(ALPACC_ERR_LIMIT, "derivation or bracket lists beyond 2^31 entries");
  const uint64_t nodes = caps ? np : np - nterm;
  // ps_b: bracket lists (u16), then the tile of every unmatched pop and slot
  // (u32, 4-byte aligned)
  const size_t off_ptile = ((size_t)(so.ob + so.ub) * 2 + 63) & ~size_t(63);
  if ((r = c.ps_b.reserve(off_ptile + (size_t)(so.ub + so.us) * 4 + 64)) ||
      (r = c.ps_c.reserve((size_t)(so.os + so.us) * 8 + 64)) ||
      (tree && ((r = c.tr_prod.reserve(nodes * g->production_bytes + 16)) || (r = c.tr_parent.reserve(nodes * 8 + 16)) ||
                (r = c.tr_tokpar.reserve(nterm * 8 + 16)))) ||
      (!tree && (r = c.ps_d.reserve(np * g->production_bytes + 16))))
    return r;
  uint16_t* ob_list = c.ps_b.as<uint16_t>();
  uint16_t* ub_list = ob_list + so.ob;
  int64_t* os_list = c.ps_c.as<int64_t>();
  int64_t* us_list = os_list + so.os;
  uint32_t* ub_tile = reinterpret_cast<uint32_t*>(c.ps_b.as<uint8_t>() + off_ptile);
  uint32_t* us_tile = ub_tile + so.ub;
  if (ntiles) {
    TileTree tb{}, ts{};
    if (!caps) HIP_TRY(tm.begin(st, "tile_tree"));
    if ((r = build_tile_trees(mb, ms, ntiles, reinterpret_cast<int32_t*>(base + off_tb),
                              reinterpret_cast<int32_t*>(base + off_tsx), st, &tb, &ts)))
      return r;
    if (!caps) HIP_TRY(tm.end(st));
    // staging sized to the largest tile when the scan's maxima are on the
    // host (fewer LDS bytes, more workgroups per CU), else to the grammar's
    // per-position maxima
    // (at capacity: the staging the context learned from an eager run of
    // the sequence, c.cap_stage_*, else the grammar's maxima)
    const int n_st = caps ? (c.cap_stage_n > 0 && c.cap_stage_n < ncap ? c.cap_stage_n : ncap)
                          : (int)(tree ? so.tmax_node : so.tmax_slot);
    const int t_st = caps ? (c.cap_stage_t > 0 && c.cap_stage_t < tcap ? c.cap_stage_t : tcap)
                          : (tree ? (int)so.tmax_term : 0);
    if (n_st > ncap || t_st > tcap) return fail(ALPACC_ERR_LIMIT, "tile counts beyond the grammar's maxima");
    const size_t lds_r = resolve_lds(n_st, t_st, br8 ? 1 : 2);
    Resolve rs{s,         reinterpret_cast<const uint16_t*>(base + off_pid),
               info,      sb,       bcap,  scap, n_st, t_st,
               ob_list,   ub_list,   ub_tile,  us_tile,  os_list,  us_list,       c.tr_prod.as<uint8_t>(),
               c.tr_parent.as<int64_t>(), c.tr_tokpar.as<int64_t>(), tree ? nullptr : c.ps_d.as<uint8_t>(),
               tree ? 1 : 0, &sout->overflow,
               g->production_bytes == 2 ? reinterpret_cast<const uint4*>(c.d_prod16) : nullptr};
    void (*kern[2])(Resolve);  // one tile per workgroup, striding
    if (g->production_bytes == 2)
      resolve_kernels<4, 4, 8, 24, true>(br8, PT, APL, kern);
    else if (ib[1] == 3)
      resolve_kernels<2, 3, 4, 3>(br8, PT, APL, kern);
    else if (narrow)
      resolve_kernels<3, 4, 4, 4>(br8, PT, APL, kern);
    else
      resolve_kernels<4, 4, 8, 24>(br8, PT, APL, kern);
    for (auto k : kern) HIP_TRY(grant_lds(reinterpret_cast<const void*>(k), lds_b));
    if (!caps) HIP_TRY(tm.begin(st, "seg_resolve"));
    hipLaunchKernelGGL(kern[stride ? 1 : 0], dim3(g_res), dim3(RWG), lds_r, st, rs);
    if (!caps) HIP_TRY(tm.end(st));
    Far fr{s, &sout->overflow, sout, info, sb, tb, ts, ob_list, ub_list, os_list, us_list, ub_tile, us_tile,
           c.tr_parent.as<int64_t>(),
           c.tr_tokpar.as<int64_t>(), tree ? 1 : 0};
    if (!caps) HIP_TRY(tm.begin(st, "far_resolve"));
    const uint64_t pops = (uint64_t)(so.ub + (tree ? so.us : 0));
    const uint32_t g_far = (uint32_t)std::min<uint64_t>(4096, std::max<uint64_t>(1, (pops + 255) / 256));
    hipLaunchKernelGGL(far_resolve, dim3(g_far), dim3(256), 0, st, fr);
    if (!caps) HIP_TRY(tm.end(st));
  }
  hipLaunchKernelGGL(seg_finish, dim3((uint32_t)((count + 255) / 256)), dim3(256), 0, st, s, info, sb, sout,
                     tree ? 1 : 0, res);
  HIP_TRY(hipGetLastError());
  if (timed) HIP_TRY(hipEventRecord(c.ev[3], st));
  // the scan's totals (maxima, overflow) and the per-input results
  c.cp_parse_down = Context::Copy{c.h_pres.ptr, sout, res_head + count * sizeof(ParseResult), hipMemcpyDeviceToHost};
  if (!c.capturing) HIP_TRY(hipMemcpyAsync(c.h_pres.ptr, sout, res_head + count * sizeof(ParseResult), hipMemcpyDeviceToHost, st));
  if (!caps) {
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(tm.collect());
  }
  if (timed) {
    float ms_ = 0;
    HIP_TRY(hipEventElapsedTime(&ms_, c.ev[2], c.ev[3]));
    c.parse_ms += ms_;
    c.parse_launches++;
  }
  c.pb_count = count;
  c.pb_res = reinterpret_cast<const ParseResult*>(c.h_pres.as<uint8_t>() + res_head);
  return 0;
}

int parse_on_device(Context& c, const uint8_t* d_tok, uint64_t n, bool tree, hipStream_t st, uint64_t* num_prods) {
  *num_prods = 0;
  int r;
  if ((r = c.pmeta.reserve(64))) return r;
  uint64_t* h_off = c.host_small + 8;  // pinned
  h_off[0] = 0;
  h_off[1] = n;
  uint64_t* d_off = c.pmeta.as<uint64_t>();
  HIP_TRY(hipMemcpyAsync(d_off, h_off, 16, hipMemcpyHostToDevice, st));
  r = parse_batch_on_device(c, d_tok, d_off, h_off, nullptr, 1, tree, st);
  if (r < 0) return r;
  const ParseResult& o = c.pb_res[0];
  if (o.flags) return ALPACC_REJECT;
  *num_prods = o.slots;
  return ALPACC_VALID;
}

}  // namespace alpacc
