// DFA lexer for gfx950 in three launches.
//
// Semantics are those of the sequential reference lexer (backends/c/lexer.c:
// 39-97): greedy tokens with no backtracking, a boundary before byte i when
// the transition on byte i is a "producing" one, ignore tokens dropped, a kept
// token longer than length_t rejects the input, and the input is accepted iff
// the state after the last byte is accepting.
//
// A launch lexes a batch of inputs packed back to back (a single input is a
// batch of one).  Every input owns whole chunks of 128 bytes (one lane each,
// its last chunk partial) and a tile is 256 consecutive chunks, so many
// small inputs share a tile; a chunk's first byte need not be aligned (loads
// are dword-aligned and bytes are realigned in registers).  The scans
// restart at the first chunk of every input.
//   L1 lex_fold      every lane folds its 128 bytes into one endomorphism
//                    (LDS tables STEP[e][byte class]); a segmented
//                    workgroup scan (compose table in L2) gives each lane
//                    its exclusive prefix inside the tile (u16, written out)
//                    and the tile its aggregate;
//   L2 lex_tile_scan three short launches scan the tile aggregates,
//                    restarting at the tiles where an input starts;
//   L3 lex_emit      every lane starts from the DFA state
//                    ENDO_INIT[tile prefix . lane prefix] and walks its bytes
//                    through the LDS token DFA (rows of next-row address |
//                    produce | kept, and the state's terminal), keeping its
//                    boundaries and kept-token closes as 128-bit masks and
//                    the terminal of each kept token in LDS; a block scan
//                    plus a decoupled lookback over (kept-token count, last
//                    boundary) -- sum and max, so a whole wave inspects 64
//                    predecessor tiles per probe -- gives the output slot and
//                    the start of every lane's first token (never before its
//                    input's first byte); each lane then visits its kept
//                    tokens from the masks (end = the close, start = the last
//                    boundary before it), stages (terminal, start, end) in
//                    LDS by slot, and the workgroup writes the
//                    structure-of-arrays output with consecutive lanes on
//                    consecutive slots.  A lane with more kept tokens than
//                    the LDS holds terminals for re-walks its bytes instead.
#include "device.hpp"
#include "lexer.hpp"
#include "runtime.hpp"

namespace alpacc {

namespace {

// 16 bytes loaded from a dword-aligned address.
typedef uint32_t u32x4a4 __attribute__((ext_vector_type(4), aligned(4)));

constexpr int WG = LEX_WG;
constexpr int NW = WG / WAVE;

// A tile: the inputs of its first and last chunks and the first of them.
struct TileSpan {
  uint64_t tbase;   // offset of the tile's first byte in the packed inputs
  uint64_t seg_lo;  // input s_lo: [seg_lo, seg_hi), its first chunk cf
  uint64_t seg_hi;
  uint32_t s_lo, s_hi, cf;
  __device__ bool multi() const { return s_hi != s_lo; }
};

template <int IPT>
__device__ __forceinline__ TileSpan tile_span(const LexParams& p, uint32_t tile) {
  TileSpan t;
  t.s_lo = p.tile_seg[2 * tile];
  t.s_hi = p.tile_seg[2 * tile + 1];
  t.cf = p.chunk_first[t.s_lo];
  t.seg_lo = p.seg_off[t.s_lo];
  t.seg_hi = p.seg_off[t.s_lo + 1];
  t.tbase = t.seg_lo + (uint64_t)(tile * WG - t.cf) * IPT;
  return t;
}

// Chunk offsets (from the tile's first chunk) of the inputs s_lo + d that
// start inside a multi-input tile, d = 1 .. s_hi - s_lo, into s_ichunk[d]
// (0xFFFF past them).  Needs a barrier before use.
__device__ __forceinline__ void tile_inputs(const LexParams& p, const TileSpan& ts, uint32_t tile,
                                            uint16_t* s_ichunk) {
  const uint32_t d = threadIdx.x + 1, m = ts.s_hi - ts.s_lo;
  if (d < (uint32_t)WG) s_ichunk[d] = d <= m ? (uint16_t)(p.chunk_first[ts.s_lo + d] - tile * WG) : (uint16_t)0xFFFFu;
}

// The same inputs' byte offsets from the tile's first byte.
__device__ __forceinline__ void input_starts(const LexParams& p, const TileSpan& ts, uint16_t* s_istart) {
  const uint32_t d = threadIdx.x + 1, m = ts.s_hi - ts.s_lo;
  if (d < (uint32_t)WG) s_istart[d] = d <= m ? (uint16_t)(p.seg_off[ts.s_lo + d] - ts.tbase) : (uint16_t)0xFFFFu;
}

// The largest d in [0, 255] with every one of sorted[1..d] <= v (index 0 is
// never read).
__device__ __forceinline__ uint32_t count_le(const uint16_t* sorted, uint32_t v) {
  uint32_t d = 0;
#pragma unroll
  for (uint32_t s = WG / 2; s >= 1; s >>= 1) d = sorted[d + s] <= v ? d + s : d;
  return d;
}

// A lane's chunk: its input, its first byte and how many bytes it holds.
struct LaneSpan {
  uint64_t base, seg_lo, seg_hi;
  uint32_t seg, valid, shift;
  bool first;  // the input's first chunk
};

template <int IPT>
__device__ __forceinline__ LaneSpan lane_span(const LexParams& p, const TileSpan& ts, uint32_t tile,
                                              const uint16_t* s_ichunk) {
  LaneSpan l;
  const uint32_t chunk = tile * WG + threadIdx.x;
  uint32_t cf = ts.cf;
  l.seg = ts.s_lo;
  l.seg_lo = ts.seg_lo;
  l.seg_hi = ts.seg_hi;
  if (ts.multi()) {
    const uint32_t d = count_le(s_ichunk, threadIdx.x);
    if (d) {
      l.seg = ts.s_lo + d;
      cf = p.chunk_first[l.seg];
      l.seg_lo = p.seg_off[l.seg];
      l.seg_hi = p.seg_off[l.seg + 1];
    }
  }
  l.first = chunk == cf;
  l.base = l.seg_lo + (uint64_t)(chunk - cf) * IPT;
  l.valid = l.base >= l.seg_hi ? 0u : (uint32_t)min((uint64_t)IPT, l.seg_hi - l.base);
  l.shift = (uint32_t)((uintptr_t)(p.in + l.base) & 3u);
  return l;
}

// The dwords holding the lane's IPT bytes (all in its input): a[0] holds its
// first byte at byte `shift`, a[IPT / 4] is read only when the lane is
// misaligned (it holds the lane's last byte), so no byte outside the dwords
// of the lane's own bytes is touched.
template <int IPT>
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

// The dwords holding the lane's first `valid` bytes (any count up to IPT;
// none for a lane without bytes, whose base may lie past the inputs), zero
// beyond them: only dwords holding one of the input's bytes are read.
template <int IPT>
__device__ __forceinline__ void load_lane_part(const LexParams& p, uint64_t base, uint32_t shift, uint32_t valid,
                                               uint32_t* a) {
  const uint32_t* src = reinterpret_cast<const uint32_t*>(p.in + base - shift);
#pragma unroll
  for (int d = 0; d <= IPT / 4; ++d) a[d] = valid && (uint32_t)(4 * d) < valid + shift ? src[d] : 0u;
}

// Bytes 4k..4k+3 of the lane.  The empty asm makes the realignment wait
// for `dep` (the walk's state), so the 32 realigned words are formed as the
// walk reaches them instead of all at once next to the raw ones.
__device__ __forceinline__ uint32_t lane_word(const uint32_t* a, int k, uint32_t shift, uint32_t dep) {
  uint32_t lo = a[k], hi = a[k + 1];
  asm volatile("" : "+v"(lo), "+v"(hi) : "v"(dep));
  return __builtin_amdgcn_alignbyte(hi, lo, shift);
}

// Inclusive segmented scan of (f, e) over a wave, with the operator
// (f1, a) . (f2, b) = (f1 | f2, f2 ? b : a then b).
__device__ __forceinline__ void seg_comp_scan(const uint16_t* __restrict__ compose, uint32_t nendo, uint32_t& f,
                                              uint32_t& e) {
  const uint32_t lane = threadIdx.x & (WAVE - 1);
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    const uint32_t of = __shfl_up(f, d, WAVE), oe = __shfl_up(e, d, WAVE);
    if (lane >= (uint32_t)d) {
      e = f ? e : compose[oe * nendo + e];
      f |= of;
    }
  }
}

// ------------------------------------------------------------------- L1

// SC: STEP holds, for endomorphism e and byte class c, the LDS byte offset
// of the next endomorphism's row (id * nclass * 2) and the class table holds
// 2c, so a byte costs one add and two LDS loads (no multiply); the row
// offsets must fit in 16 bits and 2c in 8.  Without SC the tables hold ids
// and classes.  CH independent chains walk consecutive parts of a full lane
// interleaved, since every byte is an LDS load that depends on the last, and
// their endomorphisms are then composed through the table in L2.
template <int IPT, int CH, bool SC>
__global__ __launch_bounds__(WG) void lex_fold(LexParams p) {
  extern __shared__ __align__(16) uint8_t lds[];
  uint8_t* s_class = lds;
  uint16_t* s_step = reinterpret_cast<uint16_t*>(lds + 256);
  const uint8_t* s_stepb = lds + 256;
  __shared__ uint32_t s_wave[NW];
  __shared__ uint16_t s_ichunk[WG];
  const uint32_t tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
  const uint32_t nendo = p.nendo, nclass = p.nclass, row = SC ? 2 * nclass : 1;
  for (int i = tid; i < 256; i += WG) s_class[i] = (uint8_t)(SC ? 2 * p.byte_class[i] : p.byte_class[i]);
  for (int i = tid; i < p.nendo * p.nclass; i += WG) s_step[i] = (uint16_t)(p.step[i] * row);
  __syncthreads();
  const uint16_t* __restrict__ compose = p.compose;
  auto comp = [&](uint32_t a, uint32_t b) -> uint32_t { return compose[a * nendo + b]; };
  // one byte: the walk's state x (SC: a row offset, else an id) and byte b
  auto step = [&](uint32_t x, uint32_t b) -> uint32_t {
    if (SC) return *reinterpret_cast<const uint16_t*>(s_stepb + x + s_class[b]);
    return s_step[__umul24(x, nclass) + s_class[b]];
  };

  const uint32_t nt = *p.nt_dev;
  for (uint32_t tile = blockIdx.x; tile < nt; tile += gridDim.x) {
    const TileSpan ts = tile_span<IPT>(p, tile);
    if (ts.multi()) {
      tile_inputs(p, ts, tile, s_ichunk);
      __syncthreads();
    }
    const LaneSpan ls = lane_span<IPT>(p, ts, tile, s_ichunk);
    const uint32_t valid = ls.valid;
    uint32_t e = 0;  // identity
    // the lane's bytes fully unrolled from registers (indexed with
    // constants); in the lane holding an input's last bytes, the positions
    // past them keep the state
    if (valid == IPT) {
      uint32_t a[IPT / 4 + 1];
      load_lane<IPT>(p, ls.base, ls.shift, a);
      constexpr int KC = IPT / 4 / CH;  // words per chain
      uint32_t ec[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) ec[c] = 0;
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        uint32_t wk[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) wk[c] = __builtin_amdgcn_alignbyte(a[c * KC + k + 1], a[c * KC + k], ls.shift);
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
          for (int c = 0; c < CH; ++c) ec[c] = step(ec[c], (wk[c] >> (8 * b)) & 0xFFu);
      }
#pragma unroll
      for (int c = 0; c < CH; ++c) ec[c] /= row;
#pragma unroll
      for (int w = 1; w < CH; w <<= 1)
#pragma unroll
        for (int c = 0; c + w < CH; c += 2 * w) ec[c] = comp(ec[c], ec[c + w]);
      e = ec[0];
    } else if (valid > 0) {
      uint32_t a[IPT / 4 + 1];
      load_lane_part<IPT>(p, ls.base, ls.shift, valid, a);
#pragma unroll
      for (int k = 0; k < IPT / 4; ++k) {
        const uint32_t wk = lane_word(a, k, ls.shift, e);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const uint32_t n = step(e, (wk >> (8 * b)) & 0xFFu);
          e = (uint32_t)(4 * k + b) < valid ? n : e;
        }
      }
      e /= row;
    }
    // segmented scan of (restart, endomorphism), restarting at the first
    // chunk of every input: (f1, a) . (f2, b) = (f1 | f2, f2 ? b : a then b);
    // a tile where no input starts scans the endomorphisms alone
    const bool restarts = ts.multi() || ts.cf == tile * WG;
    uint32_t fi = ls.first, ei = e;
    if (restarts) {
      seg_comp_scan(compose, nendo, fi, ei);
    } else {
#pragma unroll
      for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t o = __shfl_up(ei, d, WAVE);
        if (lane >= (uint32_t)d) ei = comp(o, ei);
      }
    }
    if (lane == WAVE - 1) s_wave[wid] = fi << 31 | ei;
    __syncthreads();
    uint32_t wf = 0, we = 0;  // over the earlier waves
    for (uint32_t i = 0; i < wid; ++i) {
      const uint32_t w = s_wave[i], f = w >> 31, x = w & 0x7FFFFFFFu;
      we = f ? x : comp(we, x);
      wf |= f;
    }
    const uint32_t lf = restarts ? __shfl_up(fi, 1, WAVE) : 0u, le = __shfl_up(ei, 1, WAVE);
    const uint32_t xf = lane ? wf | lf : wf, xe = lane ? (lf ? le : comp(we, le)) : we;
    // the lane's prefix from the tile start, or (bit 15) from its input's
    // start when that lies inside the tile
    p.lane_prefix[(uint64_t)tile * WG + tid] = (uint16_t)(ls.first ? 0x8000u : (xf ? 0x8000u | xe : xe));
    if (tid == WG - 1) {
      const uint32_t agg = (wf | fi) << 31 | (fi ? ei : comp(we, ei));
      p.tile_endo[tile] = agg;
      if (p.stream_out && tile == nt - 1) p.stream_out[1] = agg;  // for the launch's total
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------- L2

// Exclusive scan of the tile aggregates (f, e) -- f set when an input starts
// inside the tile, e the endomorphism from the last such start (else from
// the tile start) to the tile end -- with the segmented operator, in three
// short launches (one tile per thread):
//   S1 per block of 256 tiles: the tile's prefix inside its block, with a
//      flag when a restart lies between the block start and the tile
//      (written over tile_endo), and the block aggregate;
//   S2 one workgroup: each block's carry (exclusive over the blocks);
//   S3 each tile: prefix = flagged ? local : carry then local.
constexpr int TSB = 256;  // tiles per scan block

// Exclusive segmented scan of (f, e) over a workgroup of TSB threads;
// *tot receives the inclusive aggregate of the whole group.
__device__ __forceinline__ void seg_comp_block(const uint16_t* __restrict__ compose, uint32_t nendo, uint32_t f,
                                               uint32_t e, uint32_t& xf, uint32_t& xe, uint32_t& tf, uint32_t& te) {
  __shared__ uint32_t s_f[TSB / WAVE], s_e[TSB / WAVE];
  const uint32_t lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
  uint32_t fi = f, ei = e;
  seg_comp_scan(compose, nendo, fi, ei);
  if (lane == WAVE - 1) {
    s_f[wid] = fi;
    s_e[wid] = ei;
  }
  __syncthreads();
  uint32_t wf = 0, we = 0;  // exclusive over the earlier waves
  for (uint32_t w = 0; w < wid; ++w) {
    we = s_f[w] ? s_e[w] : compose[we * nendo + s_e[w]];
    wf |= s_f[w];
  }
  uint32_t lf = __shfl_up(fi, 1, WAVE), le = __shfl_up(ei, 1, WAVE);
  if (lane == 0) lf = le = 0;
  xe = lf ? le : compose[we * nendo + le];
  xf = wf | lf;
  tf = 0;
  te = 0;
  for (uint32_t w = 0; w < TSB / WAVE; ++w) {
    te = s_f[w] ? s_e[w] : compose[te * nendo + s_e[w]];
    tf |= s_f[w];
  }
}

__global__ __launch_bounds__(TSB) void lex_tile_scan1(LexParams p) {
  const uint32_t t = blockIdx.x * TSB + threadIdx.x, nt = *p.nt_dev;
  const bool ok = t < nt;
  const uint32_t a = ok ? p.tile_endo[t] : 0u;
  uint32_t xf, xe, tf, te;
  seg_comp_block(p.compose, p.nendo, a >> 31, a & 0x7FFFFFFFu, xf, xe, tf, te);
  // the tile's prefix inside the block
  if (ok) p.tile_endo[t] = (xf ? 0x80000000u : 0u) | xe;
  if (threadIdx.x == 0) p.tile_blk[blockIdx.x] = (tf ? 0x80000000u : 0u) | te;
}

__global__ __launch_bounds__(TSB) void lex_tile_scan2(LexParams p, uint32_t nblk) {
  __shared__ uint32_t s_cf, s_ce;
  if (threadIdx.x == 0) s_cf = s_ce = 0;
  __syncthreads();
  for (uint32_t b0 = 0; b0 < nblk; b0 += TSB) {
    const uint32_t b = b0 + threadIdx.x;
    const uint32_t w = b < nblk ? p.tile_blk[b] : 0u;
    uint32_t xf, xe, tf, te;
    seg_comp_block(p.compose, p.nendo, w >> 31, w & 0x7FFFFFFFu, xf, xe, tf, te);
    const uint32_t cf = s_cf, ce = s_ce;
    if (b < nblk) p.tile_blk[nblk + b] = xf ? xe : p.compose[ce * p.nendo + xe];  // carry into block b
    __syncthreads();
    if (threadIdx.x == 0) {
      s_ce = tf ? te : p.compose[ce * p.nendo + te];
      s_cf = cf | tf;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(TSB) void lex_tile_scan3(LexParams p, uint32_t nblk) {
  const uint32_t t = blockIdx.x * TSB + threadIdx.x;
  if (t >= *p.nt_dev) return;
  const uint32_t w = p.tile_endo[t], carry = p.tile_blk[nblk + t / TSB];
  const uint32_t pre = (w >> 31) ? (w & 0x7FFFFFFFu) : p.compose[carry * p.nendo + w];
  p.tile_prefix[t] = (uint16_t)pre;
  if (p.stream_out && t == *p.nt_dev - 1) {  // the endomorphism of the launch's bytes
    const uint32_t a = p.stream_out[1];
    p.stream_out[0] = (a >> 31) ? (a & 0x7FFFFFFFu) : p.compose[pre * p.nendo + (a & 0x7FFFFFFFu)];
  }
}

// ------------------------------------------------------------------- L3

constexpr uint32_t STAGE = 2176;  // kept tokens staged in LDS per round (JSON tiles: two rounds)
// LDS row of a DFA state: 256 transitions, its terminal, and a transition
// to itself that produces nothing (the bytes past an input's end)
constexpr uint32_t ROW = 258;
constexpr uint32_t STAY2 = 2 * 257;  // offset of that transition in the row
constexpr uint32_t ROWB = ROW * 2;

// Transition entries of the LDS rows (u16).  SCALED (at most 31 states):
// bits 0..13 the byte offset of the next state's row, bit 14 produce, bit
// 15 kept; otherwise the next state in bits 0..7, produce bit 8, kept bit 9.
// Rows are addressed by their LDS address (lds0 + state * ROWB, lds0 the
// address of the dynamic LDS).
template <bool SCALED>
struct RowCode {
  static constexpr uint32_t PROD = SCALED ? 14 : 8, KEPT = SCALED ? 15 : 9;
  __device__ static uint16_t encode(uint32_t d, uint32_t lds0) {
    const uint32_t next = d & 0xFFu, pr = (d >> 8) & 1u, kp = (d >> 9) & 1u;
    return (uint16_t)((SCALED ? lds0 + next * ROWB : next) | pr << PROD | kp << KEPT);
  }
  __device__ static uint32_t next_row(uint32_t x, uint32_t lds0) {
    return SCALED ? (x & 0x3FFFu) : lds0 + (x & 0xFFu) * ROWB;
  }
};

typedef __attribute__((address_space(3))) uint8_t lds_u8;
typedef __attribute__((address_space(3))) uint16_t lds_u16;
__device__ __forceinline__ uint32_t lds_addr(const void* p) {
  return (uint32_t)(uintptr_t)(const lds_u8*)p;
}
__device__ __forceinline__ uint32_t lds_load16(uint32_t a) { return *(const lds_u16*)a; }
__device__ __forceinline__ void lds_store8(uint32_t a, uint32_t v) { *(lds_u8*)a = (uint8_t)v; }

// DFA state of the lane's first byte.
// DFA state of the lane's first byte (after the bytes of earlier launches
// of a stream, carry_endo, when there are any).
__device__ __forceinline__ uint32_t lane_state(const LexParams& p, uint32_t tile, uint32_t tid) {
  const uint32_t lp = p.lane_prefix[(uint64_t)tile * WG + tid];
  uint32_t pre = (lp & 0x8000u) ? lp & 0x7FFFu : p.compose[(uint32_t)p.tile_prefix[tile] * p.nendo + lp];
  if (p.carry_endo) pre = p.compose[p.carry_endo * p.nendo + pre];
  return p.endo_init[pre];
}

// Workgroup words of lex_emit, kept after its dynamic LDS arrays (the
// kernel has no static LDS, so the DFA rows start at LDS address 0 and a
// row offset is an LDS address).
struct EmitWords {
  uint32_t tile, pre_cnt, pre_last, pad;
  uint32_t wave_cnt[NW], wave_last[NW];
};

// LDS layout of lex_emit: DFA rows, the staged slots (a word of start |
// end << 16, both tile-relative, and a half-word of terminal | input << 8,
// the input counted from the tile's first), the lanes' kept-token
// terminals (rcap rows of WG bytes and four sink rows), the workgroup
// words.
struct EmitLayout {
  uint32_t se, ts, lterm, words, total;
  __host__ __device__ EmitLayout(uint32_t nstate, uint32_t rcap) {
    se = (nstate * ROWB + 15u) & ~15u;
    ts = se + 4 * STAGE;
    lterm = ts + 2 * STAGE;
    words = (lterm + (rcap + 4) * WG + 15u) & ~15u;
    total = words + (uint32_t)sizeof(EmitWords);
  }
};

template <bool SCALED, int IPT>
__global__ __launch_bounds__(WG) void lex_emit(LexParams p) {
  using RC = RowCode<SCALED>;
  extern __shared__ __align__(16) uint8_t lds[];
  const uint32_t rcap = p.rcap;
  const EmitLayout L((uint32_t)p.nstate, rcap);
  uint16_t* s_rows = reinterpret_cast<uint16_t*>(lds);              // [nstate][ROW]
  uint32_t* s_se = reinterpret_cast<uint32_t*>(lds + L.se);  // [STAGE] start | end << 16
  uint16_t* s_ts = reinterpret_cast<uint16_t*>(lds + L.ts);  // [STAGE] terminal | input << 8
  uint16_t* s_ichunk = s_ts;                                  // [WG] before the staging
  // [WG] after walk 1, in its sink rows
  uint16_t* s_istart = reinterpret_cast<uint16_t*>(lds + L.lterm + rcap * WG);
  EmitWords& w = *reinterpret_cast<EmitWords*>(lds + L.words);
  const uint32_t tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
  const uint32_t lds0 = lds_addr(lds);
  for (uint32_t i = tid; i < (uint32_t)p.nstate * ROW; i += WG) {
    const uint32_t s = i / ROW, b = i - s * ROW;
    s_rows[i] = b == 256 ? (uint16_t)p.token[s] : RC::encode(b == 257 ? s : p.delta[s * 256 + b], lds0);
  }
  const int ignore = p.ignore;
  // s_lterm[r][lane] is the byte at LDS address lterm0 + r * WG + swz, swz
  // = 4 * (lane % 64) + lane / 64: the 64 lanes of a wave write 64
  // different banks whatever their r (a plain lane-order byte would put
  // four lanes on each bank).  A lane's writes past rcap tokens go to its
  // own byte of the sink rows.
  const uint32_t swz = 4u * (tid & (WAVE - 1)) + tid / WAVE;
  const uint32_t lterm0 = lds0 + L.lterm, sink = lterm0 + rcap * WG + swz;
  const uint32_t nt = *p.nt_dev;

  for (;;) {
    if (tid == 0) w.tile = atomicAdd(p.ticket, 1u);
    __syncthreads();
    const uint32_t tile = __builtin_amdgcn_readfirstlane(w.tile);
    if (tile >= nt) break;
    const TileSpan ts = tile_span<IPT>(p, tile);
    if (ts.multi()) {
      tile_inputs(p, ts, tile, s_ichunk);
      __syncthreads();
    }
    const LaneSpan ls = lane_span<IPT>(p, ts, tile, s_ichunk);
    const uint64_t tbase = ts.tbase, base = ls.base;
    const uint32_t valid = ls.valid;
    const uint32_t row0 = lds0 + lane_state(p, tile, tid) * ROWB;

    // walk 1: the lane's boundaries (bm) and kept-token closes (km) as bit
    // masks, the terminal of its r-th kept token in s_lterm[r][lane] (every
    // byte writes the slot of the next kept token).  The lane's bytes are
    // walked fully unrolled from registers, with the write address clamped
    // to the sink once per 4 bytes (so at most 3 rows past it); in the lane
    // holding an input's last bytes, the positions past them take the row's
    // stay transition.
    uint32_t row = row0, waddr = lterm0 + swz, bm[IPT / 32] = {}, km[IPT / 32] = {};
    // off2 = 2 * byte: the transition's offset in the row
    auto step1 = [&](uint32_t off2, uint32_t& mb, uint32_t& mk, uint32_t bit) {
      const uint32_t x = lds_load16(row + off2);
      const uint32_t t = lds_load16(row + 512);
      const uint32_t kp = (x >> RC::KEPT) & 1u;
      mb |= ((x >> RC::PROD) & 1u) << bit;
      mk |= kp << bit;
      lds_store8(waddr, t);
      waddr += kp << 8;
      asm volatile("" : "+v"(waddr));  // a running address, not rebuilt from a count
      row = RC::next_row(x, lds0);
    };
    // One walk a wave: the plain one when every lane of the wave is full,
    // else the bounded one in every lane (a wave taking both would walk its
    // bytes twice)
    if (__ballot(valid != IPT) == 0) {
      uint32_t a[IPT / 4 + 1];
      load_lane<IPT>(p, base, ls.shift, a);
#pragma unroll
      for (int k = 0; k < IPT / 4; ++k) {
        const uint32_t wk = lane_word(a, k, ls.shift, row);
        // doubled bytes 0 and 2, 1 and 3 in the halves of two words
        uint32_t e02 = (wk << 1) & 0x01FE01FEu, e13 = (wk >> 7) & 0x01FE01FEu;
        asm volatile("" : "+v"(e02), "+v"(e13));  // kept paired, not re-extracted per byte
        waddr = waddr < sink ? waddr : sink;
        step1(e02 & 0xFFFFu, bm[k >> 3], km[k >> 3], (4 * k) & 31);
        step1(e13 & 0xFFFFu, bm[k >> 3], km[k >> 3], (4 * k + 1) & 31);
        step1(e02 >> 16, bm[k >> 3], km[k >> 3], (4 * k + 2) & 31);
        step1(e13 >> 16, bm[k >> 3], km[k >> 3], (4 * k + 3) & 31);
      }
    } else {
      uint32_t a[IPT / 4 + 1];
      load_lane_part<IPT>(p, base, ls.shift, valid, a);
#pragma unroll
      for (int k = 0; k < IPT / 4; ++k) {
        const uint32_t wk = lane_word(a, k, ls.shift, row);
        waddr = waddr < sink ? waddr : sink;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const uint32_t j = 4 * k + b;
          step1(j < valid ? ((wk >> (8 * b)) & 0xFFu) * 2u : STAY2, bm[k >> 3], km[k >> 3], j & 31);
        }
      }
    }
    uint32_t cnt = 0;  // kept tokens closed inside the lane
#pragma unroll
    for (int q = 0; q < IPT / 32; ++q) cnt += __builtin_popcount(km[q]);
    // more kept tokens than walk 1 kept terminals for: this lane re-walks
    const bool lane_over = cnt > rcap;
    // the input's last byte (a launch that a stream continues has none)
    const bool owns_end = valid > 0 && base + valid == ls.seg_hi && p.is_final;
    uint32_t fin_t = 0;
    bool fin_kept = false;
    if (owns_end) {
      fin_t = lds_load16(row + 512);
      if ((int)fin_t == p.dead) {
        atomicOr(p.seg_flags + ls.seg, LEX_REJECT_FINAL);
      } else if ((int)fin_t != ignore) {
        fin_kept = true;
      }
    }
    cnt += fin_kept;
    uint32_t last = 0;
#pragma unroll
    for (int q = 0; q < IPT / 32; ++q)
      if (bm[q]) last = (uint32_t)base + 32 * q + 32 - __builtin_clz(bm[q]);  // position + 1: 0 = none

    // block scan of (count, last boundary)
    uint32_t icnt = cnt, ilast = last;
    wave_scan_add_max(icnt, ilast);
    if (lane == WAVE - 1) {
      w.wave_cnt[wid] = icnt;
      w.wave_last[wid] = ilast;
    }
    __syncthreads();
    // walk 1 is over, its sink rows free until the next tile: the byte
    // offsets of the inputs starting in the tile, read from the first
    // round's write-out on
    if (ts.multi()) input_starts(p, ts, s_istart);
    uint32_t wc = 0, wl = 0, tc = 0, tl = 0;
    for (uint32_t i = 0; i < NW; ++i) {
      if (i < wid) {
        wc += w.wave_cnt[i];
        wl = wl > w.wave_last[i] ? wl : w.wave_last[i];
      }
      tc += w.wave_cnt[i];
      tl = tl > w.wave_last[i] ? tl : w.wave_last[i];
    }
    const uint32_t pc = __shfl_up(icnt, 1, WAVE), pl = __shfl_up(ilast, 1, WAVE);
    const uint32_t ecnt = wc + (lane ? pc : 0u);
    const uint32_t elast = lane ? (wl > pl ? wl : pl) : wl;

    // decoupled lookback by the last wave (its lanes' tokens mostly fall in
    // a later round), 64 predecessor tiles per probe, while the other waves
    // stage the first round; its results are read after the round's barrier
    if (wid == NW - 1) {
      const uint64_t agg = (uint64_t)tc | ((uint64_t)tl << 31);
      uint32_t xc = 0, xl = 0;
      if (tile == 0) {
        if (lane == 0) publish64(p.count_status, ST64_PRE | agg);
      } else {
        if (lane == 0) publish64(p.count_status + tile, ST64_AGG | agg);
        int64_t top = (int64_t)tile - 1;
        const uint32_t win = p.lb_win;  // predecessor tiles per probe, 1..WAVE
        for (;;) {
          const int64_t idx = top - (int64_t)lane;
          // lanes past the window read nothing: an aggregate of zero
          const uint64_t st = lane >= win ? ST64_AGG : (idx >= 0 ? observe64(p.count_status + idx) : ST64_PRE);
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
          top -= win;
        }
        if (lane == 0) {
          const uint32_t ic = xc + tc, il = xl > tl ? xl : tl;
          publish64(p.count_status + tile, ST64_PRE | (uint64_t)ic | ((uint64_t)il << 31));
        }
      }
      if (lane == 0) {
        w.pre_cnt = xc;
        w.pre_last = xl;
      }
    }
    // where the lane's first token starts: the last boundary of an earlier
    // lane, else (NO_START) the last boundary before the tile, known after
    // the lookback; never before the token's input's first byte (applied at
    // the write-out, where a boundary of an earlier input gives way to it).
    // Boundaries travel as position + 1, 0 meaning none; staged starts are
    // tile-relative.  A token belongs to the input of the lane that closes
    // it: a lane closing a token at its first byte is not its input's first
    // chunk (no token closes at an input's first byte), so the token's last
    // byte lies in the same input.
    constexpr uint32_t NO_START = 0xFFFFu;
    const uint32_t start_in = elast ? elast - 1 - (uint32_t)tbase : NO_START;
    const uint32_t dl8 = (ls.seg - ts.s_lo) << 8;  // the lane's input, from the tile's first
    uint64_t slot0 = 0;
    uint32_t carried = 0;
    bool overflow = false;
    const uint32_t lane_off = (uint32_t)(base - tbase);  // the tile's bytes are contiguous

    // per round of STAGE slots, the lanes owning slots of the round stage
    // (terminal, start, end) by slot; then the workgroup writes the round
    // out with consecutive lanes on consecutive slots
    uint32_t lo = 0;
    do {
      const uint32_t hi = lo + STAGE;
      if (cnt > 0 && ecnt < hi && ecnt + cnt > lo) {
        const uint32_t d = ecnt - lo;  // staging index of the lane's first token (mod 2^32)
        uint32_t k = 0, start = start_in;
        if (!lane_over) {
          // kept token k closes before byte j (bit j of km) and starts at the
          // lane's last boundary before j, else where the lane's first token starts
#pragma unroll
          for (int q = 0; q < IPT / 32; ++q) {
            uint32_t m = km[q];
            const uint32_t pm = bm[q], qb = lane_off + 32 * q;
            while (m) {
              const uint32_t jb = __builtin_ctz(m);
              m &= m - 1;
              const uint32_t below = pm & ((1u << jb) - 1u);
              const uint32_t i = d + k;
              if (i < STAGE) {
                s_ts[i] = (uint16_t)(lds[L.lterm + k * WG + swz] | dl8);
                s_se[i] = (below ? qb + 31 - __builtin_clz(below) : start) | (qb + jb) << 16;
              }
              ++k;
            }
            start = pm ? qb + 31 - __builtin_clz(pm) : start;
          }
        } else {
          // the lane closes more kept tokens than walk 1 kept terminals
          // for: re-walk its bytes through the DFA, four per dword load
          // (only the dwords holding its bytes, as in walk 1)
          uint32_t r = row0;
          const uint32_t* src = reinterpret_cast<const uint32_t*>(p.in + base - ls.shift);
          for (uint32_t k4 = 0; 4 * k4 < valid; ++k4) {
            const uint32_t lo_w = src[k4], hi_w = 4 * k4 + 4 < valid + ls.shift ? src[k4 + 1] : 0u;
            const uint32_t wk = __builtin_amdgcn_alignbyte(hi_w, lo_w, ls.shift);
            for (uint32_t b = 0; b < 4 && 4 * k4 + b < valid; ++b) {
              const uint32_t j = 4 * k4 + b;
              const uint32_t x = lds_load16(r + 2 * ((wk >> (8 * b)) & 0xFFu));
              if ((x >> RC::KEPT) & 1u) {
                const uint32_t i = d + k;
                if (i < STAGE) {
                  s_ts[i] = (uint16_t)((lds_load16(r + 512) & 0xFFu) | dl8);
                  s_se[i] = start | (lane_off + j) << 16;
                }
                ++k;
              }
              start = ((x >> RC::PROD) & 1u) ? lane_off + j : start;
              r = RC::next_row(x, lds0);
            }
          }
        }
        if (fin_kept) {
          const uint32_t i = d + k;
          if (i < STAGE) {
            s_ts[i] = (uint16_t)((fin_t & 0xFFu) | dl8);
            s_se[i] = start | (lane_off + valid) << 16;
          }
        }
      }
      __syncthreads();
      if (lo == 0) {
        slot0 = w.pre_cnt;     // output slot of the tile's first kept token
        carried = w.pre_last;  // the last boundary before the tile
        if (ls.first) p.seg_tok[ls.seg] = slot0 + ecnt;
        if (tid == 0 && tile == nt - 1) {
          p.seg_tok[p.nseg] = slot0 + tc;
          if (p.stream_out) p.stream_out[2] = carried > tl ? carried : tl;  // the launch's last boundary + 1
        }
      }
      const uint32_t m = tc - lo < STAGE ? tc - lo : STAGE;
      const uint64_t max_len = p.max_len;
      if (!ts.multi()) {
        // one input: a carried start never lies before its first byte; with
        // no boundary since it, the token started there or, in a launch that
        // continues a stream, where the stream's open token started
        const int64_t lo_in = (int64_t)ts.seg_lo;
        const int64_t cs = carried ? ((int64_t)carried - 1 > lo_in ? (int64_t)carried - 1 : lo_in)
                                   : lo_in + p.carry_start;
        for (uint32_t i = tid; i < m; i += WG) {
          const uint64_t slot = slot0 + lo + i;
          const uint32_t se = s_se[i], b0 = se & 0xFFFFu;
          const int64_t b = b0 == NO_START ? cs : (int64_t)tbase + b0;
          const uint64_t len = (uint64_t)((int64_t)tbase + (se >> 16) - b);
          overflow |= len > max_len;
          p.terminals[slot] = (uint8_t)s_ts[i];
          p.starts[slot] = b - lo_in + p.start_bias;
          store_length(p, slot, len);
        }
      } else {
        for (uint32_t i = tid; i < m; i += WG) {
          const uint64_t slot = slot0 + lo + i;
          const uint32_t se = s_se[i], e = se >> 16, tw = s_ts[i];
          // the token's input, staged with it
          const uint32_t d = tw >> 8;
          const uint32_t seg = ts.s_lo + d;
          const uint64_t seg_lo = d ? tbase + s_istart[d] : ts.seg_lo;
          const uint32_t b0 = se & 0xFFFFu;
          uint64_t b = b0 == NO_START ? (carried ? carried - 1 : 0u) : tbase + b0;
          b = b > seg_lo ? b : seg_lo;
          const uint64_t len = tbase + e - b;
          if (len > max_len) atomicOr(p.seg_flags + seg, LEX_REJECT_OVERFLOW);
          p.terminals[slot] = (uint8_t)tw;
          p.starts[slot] = (int64_t)(b - seg_lo);
          store_length(p, slot, len);
        }
      }
      __syncthreads();
      lo += STAGE;
    } while (lo < tc);
    if (overflow) atomicOr(p.seg_flags + ts.s_lo, LEX_REJECT_OVERFLOW);
  }
}

// Persistent grid of a kernel: its resident workgroups per CU (at most wpc
// when wpc > 0) times the CUs.
template <typename K>
hipError_t occupancy_grid(K k, size_t lds, int wpc, int* grid) {
  int dev = 0, cus = 0, per = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  e = grant_lds(reinterpret_cast<const void*>(k), lds);
  if (e != hipSuccess) return e;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(k), WG, lds);
  if (e != hipSuccess) return e;
  if (wpc > 0 && per > wpc) per = wpc;
  *grid = cus * (per < 1 ? 1 : per);
  return hipSuccess;
}

}  // namespace

size_t lex_fold_lds(const LexParams& p) { return 256 + (size_t)p.nendo * p.nclass * 2; }
namespace {
// Rows hold next-row LDS addresses when those fit in 14 bits: lex_emit has
// no static LDS, so its dynamic LDS starts at address 0.
bool emit_scaled(const LexParams& p) {
  static const bool no_static_lds = [] {
    hipFuncAttributes fa{}, fb{};
    return hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(lex_emit<true, 128>)) == hipSuccess &&
           hipFuncGetAttributes(&fb, reinterpret_cast<const void*>(lex_emit<true, 64>)) == hipSuccess &&
           fa.sharedSizeBytes == 0 && fb.sharedSizeBytes == 0;
  }();
  return no_static_lds && (uint32_t)p.nstate * ROWB <= 0x4000u;
}
// Kept tokens per lane whose terminals walk 1 keeps in LDS: as many as fit
// in 40 KiB per workgroup (four workgroups per CU), between 16 and the
// lane's bytes.
uint32_t emit_rcap(const LexParams& p) {
  const uint32_t budget = 40 * 1024, fixed = EmitLayout((uint32_t)p.nstate, 0).total + 16;
  const uint32_t r = fixed < budget ? (budget - fixed) / WG : 0, ipt = (uint32_t)p.ipt;
  return r < 16 ? 16 : (r > ipt ? ipt : r);
}

// The fold's tables hold LDS row offsets when those fit in 16 bits and
// doubled classes in 8.
bool fold_scaled(const LexParams& p) { return (uint64_t)p.nendo * p.nclass * 2 <= 0x10000u && p.nclass <= 128; }

// The kernels of an items-per-thread shape.
struct LexKernels {
  void (*fold)(LexParams);
  void (*emit)(LexParams);
};
LexKernels lex_kernels(const LexParams& p) {
  const bool sc = emit_scaled(p);
  const bool fs = fold_scaled(p);
  if (p.ipt == 64) return {fs ? lex_fold<64, 2, true> : lex_fold<64, 1, false>, sc ? lex_emit<true, 64> : lex_emit<false, 64>};
  return {fs ? lex_fold<128, 2, true> : lex_fold<128, 1, false>, sc ? lex_emit<true, 128> : lex_emit<false, 128>};
}
}  // namespace

size_t lex_emit_lds(const LexParams& p) { return EmitLayout((uint32_t)p.nstate, emit_rcap(p)).total; }
namespace {
// tile_seg[2t] and tile_seg[2t + 1]: the inputs of the first and the last
// chunk of lexer tile t (the last i with chunk_first[i] <= chunk).
__global__ void lex_tile_map(const uint32_t* chunk_first, uint32_t nseg, const uint32_t* nt_dev, uint32_t* tile_seg) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * *nt_dev) return;
  const uint32_t t = i >> 1, nchunks = chunk_first[nseg];
  const uint32_t c = (i & 1) ? min(t * WG + WG - 1, nchunks - 1) : t * WG;
  uint32_t lo = 0, hi = nseg;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (chunk_first[mid] <= c)
      lo = mid;
    else
      hi = mid;
  }
  tile_seg[i] = lo;
}
}  // namespace

hipError_t lex_grid_size(const LexParams& p, int wpc, int* fold_grid, int* emit_grid) {
  const LexKernels k = lex_kernels(p);
  hipError_t e = occupancy_grid(k.fold, lex_fold_lds(p), wpc, fold_grid);
  if (e != hipSuccess) return e;
  return occupancy_grid(k.emit, lex_emit_lds(p), wpc, emit_grid);
}

bool lex_shape_ok(int ipt, int lb_win) { return (ipt == 64 || ipt == 128) && lb_win >= 1 && lb_win <= WAVE; }

hipError_t lex_launch(const LexParams& p, int fold_grid, int emit_grid, hipStream_t st, KernelTimer& tm) {
  const uint32_t nt = p.ntiles;
  int g1 = fold_grid < (int)nt ? fold_grid : (int)nt;
  int g3 = emit_grid < (int)nt ? emit_grid : (int)nt;
  hipError_t e;
  hipLaunchKernelGGL(lex_tile_map, dim3((2 * nt + 255) / 256), dim3(256), 0, st, p.chunk_first, p.nseg, p.nt_dev,
                     p.tile_seg);
  const LexKernels k = lex_kernels(p);
  if ((e = tm.begin(st, "lex_fold")) != hipSuccess) return e;
  hipLaunchKernelGGL(k.fold, dim3(g1), dim3(WG), lex_fold_lds(p), st, p);
  if ((e = tm.end(st)) != hipSuccess || (e = tm.begin(st, "lex_tile_scan")) != hipSuccess) return e;
  const uint32_t nblk = (nt + TSB - 1) / TSB;
  hipLaunchKernelGGL(lex_tile_scan1, dim3(nblk), dim3(TSB), 0, st, p);
  hipLaunchKernelGGL(lex_tile_scan2, dim3(1), dim3(TSB), 0, st, p, nblk);
  hipLaunchKernelGGL(lex_tile_scan3, dim3(nblk), dim3(TSB), 0, st, p, nblk);
  if ((e = tm.end(st)) != hipSuccess || (e = tm.begin(st, "lex_emit")) != hipSuccess) return e;
  LexParams q = p;
  q.rcap = emit_rcap(p);
  hipLaunchKernelGGL(k.emit, dim3(g3), dim3(WG), lex_emit_lds(p), st, q);
  if ((e = tm.end(st)) != hipSuccess) return e;
  return hipGetLastError();
}

}  // namespace alpacc
