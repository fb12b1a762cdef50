// LLP parser pipeline on the device (parse_seg.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "runtime.hpp"

namespace alpacc {

// Per-input result of a parser launch.  flags == 0: valid.  Slots are the
// full derivation (every production, flat order); nodes are the compact
// tree's non-terminal productions; terminal productions ("term") index the
// token parents.  Offsets are into the launch's output arrays.
struct ParseResult {
  uint32_t flags, pad;
  uint64_t slot_off, slots;
  uint64_t node_off, nodes;
  uint64_t term_off;
};

// Parse `count` strings of terminal ids (combined numbering, u8) packed in
// device memory at d_tok, string i being ids [tok_off[i], tok_off[i + 1])
// (d_tok_off on the device, h_tok_off its host copy).  Strings with
// h_skip[i] != 0 are not parsed (h_skip may be null).  With `tree` the
// compact trees land in c.tr_* (parents relative to their input), else the
// flat derivations in c.ps_d.  Returns 0 or an error; per-input results in
// c.pb_res[count].
//
// With caps (caps->bytes bounding the tokens of all strings), nothing is
// read back before the sequence ends: the tiles are planned on the device
// from d_tok_off, the device flags d_skip and *d_live (strings at or past it
// are padding), h_tok_off and h_skip are not read, and c.pb_res is written
// when the stream reaches its copy.
int parse_batch_on_device(Context& c, const uint8_t* d_tok, const uint64_t* d_tok_off, const uint64_t* h_tok_off,
                          const uint32_t* h_skip, uint64_t count, bool tree, hipStream_t st,
                          const BatchCaps* caps = nullptr, const uint32_t* d_skip = nullptr,
                          const uint32_t* d_live = nullptr);

// Bound on how far one framed position moves any depth or list offset of
// the parser's scans: the largest of a window's bracket / slot pushes and
// pops.  Positions times this must stay below 2^31 in one launch.
int parse_per_position(const GrammarDesc* g);
// Framed positions one parser launch takes for sure (any count of inputs
// up to this many): positions * per_position + inputs < 2^31.
uint64_t parse_max_positions(const GrammarDesc* g);

// The parser's tile shapes: positions per tile (512, 1024, 2048) and per
// lane of the resolve pass (4; or 2, 4, 8 at 1024 positions per tile).
bool parse_shape_ok(int positions_per_tile, int positions_per_lane);

// After the copy c.cp_parse_down of a launch has landed in host memory:
// whether seg_resolve met a tile beyond its staging (the launch's results
// are not valid), and the largest per-tile node and terminal-production
// counts.
bool parse_scan_overflow(const Context& c);
void parse_scan_maxima(const Context& c, bool tree, int* nodes, int* terms);

// Largest bracket symbol of the grammar's pair records.
uint32_t bracket_symbol_max(const GrammarDesc* g);

// One string of n ids: returns 1 (valid) / 0 (rejected) / < 0 (error) and
// the derivation length in *num_prods (trees: nodes = *num_prods - n).
int parse_on_device(Context& c, const uint8_t* d_tok, uint64_t n, bool tree, hipStream_t st, uint64_t* num_prods);

}  // namespace alpacc
