// LLP parser pipeline on the device (parser.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "runtime.hpp"

namespace alpacc {

// Parse n terminal ids (u8, device) with the combined-mode numbering.
// Returns 1 (valid) / 0 (rejected) / < 0 (error) and the derivation length in
// *num_prods.  With `tree` the parents and the compact productions-only tree
// land in c.tr_* (num_nodes = *num_prods - n); without it the flat
// derivation (u8 productions) lands in c.ps_d.
int parse_on_device(Context& c, const uint8_t* d_tok, uint64_t n, bool tree, hipStream_t st, uint64_t* num_prods);

}  // namespace alpacc
