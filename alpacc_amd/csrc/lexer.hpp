// Parameters and entry points of the DFA lexer (lexer.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "timer.hpp"

namespace alpacc {

constexpr int LEX_WG = 256;   // lanes per workgroup (4 waves)
constexpr int LEX_IPT = 128;  // input bytes per lane (a chunk) by default; 64 is the other shape
constexpr int LEX_LB_WIN = 16;  // predecessor tiles per lookback probe by default

// Input bytes of one lexer launch: positions and counts travel in 31-bit
// fields of the lookback words.
constexpr uint64_t LEX_MAX_LAUNCH_BYTES = (1ull << 31) - 1;
// Chunk of an input beyond one launch, lexed through the stream path.
constexpr uint64_t LEX_STREAM_CHUNK = 1ull << 30;

constexpr uint32_t LEX_REJECT_FINAL = 1u;     // last state not accepting
constexpr uint32_t LEX_REJECT_OVERFLOW = 2u;  // a kept token is longer than length_t

// A launch lexes `nseg` inputs packed back to back in `in` (a single input
// is a batch of one).  Every input owns whole chunks of LEX_IPT bytes (its
// last one partial) -- at least one, so an empty input has a chunk with no
// bytes -- and chunk_first[i] is the first chunk of input i
// (chunk_first[nseg] = all chunks).  Tiles are LEX_WG consecutive chunks, so
// a tile holds up to LEX_WG inputs and its bytes are one contiguous range.
struct LexParams {
  const uint8_t* in;
  uint64_t in_bytes;            // bytes of the packed inputs at `in`
  const uint64_t* seg_off;      // [nseg + 1] byte offset of each input in `in`
  const uint32_t* chunk_first;  // [nseg + 1]
  uint32_t* tile_seg;           // [2 * ntiles] inputs of each tile's first and last chunks
  uint32_t nseg;
  // tables
  const uint8_t* byte_class;
  const uint16_t* step;
  const uint16_t* compose;
  const uint8_t* endo_init;
  const uint16_t* delta;  // next | produce<<8 | kept<<9
  const uint8_t* token;   // [nstate] terminal each state accepts
  int nstate, nendo, nclass;
  int ignore, dead;
  uint64_t max_len;
  int length_bytes;
  // outputs (capacity: the total input bytes); starts relative to the input
  uint8_t* terminals;
  int64_t* starts;
  void* lengths;
  uint32_t* seg_flags;  // [nseg] LEX_REJECT_* bits, zeroed before the launch
  uint64_t* seg_tok;    // [nseg + 1] first token slot of each input; [nseg] = all tokens
  // scan state
  // [ntiles * LEX_WG] exclusive endomorphism of each lane inside its tile,
  // bit 15 set when its input starts inside the tile (at or before the lane)
  uint16_t* lane_prefix;
  uint32_t* tile_endo;     // [ntiles] tile aggregates, bit 31 set when an input starts inside the tile
  uint16_t* tile_prefix;   // [ntiles] exclusive tile prefixes inside the input
  uint32_t* tile_blk;      // [2 * ceil(ntiles / 256)] scan blocks' aggregates, then their carries
  uint64_t* count_status;  // [ntiles] lookback words, zeroed before every launch
  uint32_t* ticket;        // [1] zeroed before every launch
  // tiles of the launch: *nt_dev in device memory; ntiles is an upper bound
  // (equal to it unless the launch sequence is sized for a batch capacity)
  // that sizes the grids and the scan state
  const uint32_t* nt_dev;
  uint32_t ntiles;
  uint32_t rcap;  // set by lex_launch: kept tokens per lane lex_emit keeps terminals for
  // shape: input bytes per lane (64 or 128: the chunk) and predecessor
  // tiles per lookback probe (1..64)
  int ipt, lb_win;
  // a launch of one input that continues a stream (zero / one / null
  // otherwise): carry_endo = the endomorphism of the stream's earlier bytes,
  // carry_start = where its open token started (relative to this launch,
  // <= 0), start_bias added to every start written, is_final = the stream
  // ends with this launch (its last token and the final-state check),
  // stream_out[0] = the endomorphism of this launch's bytes, [2] = its last
  // boundary + 1 (0: none), [1] scratch
  uint32_t carry_endo;
  int64_t carry_start, start_bias;
  int is_final;
  uint32_t* stream_out;
};

__device__ __forceinline__ void store_length(const LexParams& p, uint64_t slot, uint64_t len) {
  switch (p.length_bytes) {
    case 1: static_cast<uint8_t*>(p.lengths)[slot] = (uint8_t)len; break;
    case 2: static_cast<uint16_t*>(p.lengths)[slot] = (uint16_t)len; break;
    case 4: static_cast<uint32_t*>(p.lengths)[slot] = (uint32_t)len; break;
    default: static_cast<uint64_t*>(p.lengths)[slot] = len; break;
  }
}

size_t lex_fold_lds(const LexParams& p);
size_t lex_emit_lds(const LexParams& p);
// Persistent grids of the fold and emit kernels of p's shape, at most wpc
// workgroups per CU (0: as many as fit).
hipError_t lex_grid_size(const LexParams& p, int wpc, int* fold_grid, int* emit_grid);
bool lex_shape_ok(int ipt, int lb_win);
hipError_t lex_launch(const LexParams& p, int fold_grid, int emit_grid, hipStream_t st, KernelTimer& tm);

}  // namespace alpacc
