// LLP parser pipeline on the device (parser.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "runtime.hpp"

namespace alpacc {

// Parse n terminal ids (u8, device) with the combined-mode numbering.
// On success: returns 1 (valid) / 0 (rejected); the derivation (u8
// productions) is in c.ps_b with *num_prods entries.  When `tree` is set it
// also computes the parents and the compacted productions-only tree into
// c.tr_* (num_nodes = *num_prods - n).
int parse_on_device(Context& c, const uint8_t* d_tok, uint64_t n, bool tree, hipStream_t st, uint64_t* num_prods);

}  // namespace alpacc
