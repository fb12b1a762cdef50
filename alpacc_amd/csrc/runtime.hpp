// Host-side context of one resident grammar (shared by the runtime sources).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "grammar.hpp"

namespace alpacc {

// Grow-only device allocation: steady-state calls allocate nothing.
struct DevBuf {
  void* ptr = nullptr;
  size_t cap = 0;
  int reserve(size_t bytes);
  void release();
  template <typename T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

struct Context {
  const GrammarDesc* g = nullptr;
  int device = 0;
  hipStream_t stream = nullptr;
  int lex_grid = 0, lex_grid2 = 0;  // persistent grids of the fold and emit kernels
  size_t resolve_lds_set = 0;        // dynamic LDS granted to the parser's resolve kernel
  // resident tables
  uint8_t* d_class = nullptr;
  uint16_t* d_step = nullptr;
  uint16_t* d_compose = nullptr;
  uint8_t* d_endo_init = nullptr;
  uint16_t* d_delta_lex = nullptr;
  uint16_t* d_delta_both = nullptr;
  uint16_t* d_window = nullptr;
  int32_t* d_spans = nullptr;
  uint8_t* d_profile = nullptr;
  uint8_t* d_pair_rec = nullptr;
  uint8_t* d_prod_info = nullptr;
  uint16_t* d_stacks = nullptr;
  uint8_t* d_productions = nullptr;
  uint8_t* d_arity = nullptr;
  uint8_t* d_isterm = nullptr;
  // workspaces
  DevBuf small;                       // counters, flags, tickets
  DevBuf lx_term, lx_start, lx_len;   // lexer outputs
  DevBuf lx_status;                   // lookback words of the lexer
  DevBuf in_copy;                     // staging of host inputs
  DevBuf ps_a, ps_b, ps_c, ps_d, ps_e, ps_f, ps_g, ps_h, ps_i;  // parser temporaries
  DevBuf tr_prod, tr_parent, tr_tokpar;                         // tree outputs
  uint64_t* host_small = nullptr;    // pinned: count / flags read-back
  // optional per-kernel timing with events on the launch stream
  bool timing = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double lex_ms = 0, parse_ms = 0;
  uint64_t lex_launches = 0, parse_launches = 0;

  std::vector<DevBuf*> all_buffers() {
    return {&small, &lx_term, &lx_start, &lx_len, &lx_status, &in_copy, &ps_a, &ps_b, &ps_c,
            &ps_d,  &ps_e,    &ps_f,     &ps_g,   &ps_h,      &ps_i,    &tr_prod, &tr_parent, &tr_tokpar};
  }
};

// Opaque handle of the C ABI.
}  // namespace alpacc
struct alpacc_hip_ctx {
  alpacc::Context c;
};
namespace alpacc {

int fail(int code, const char* fmt, ...);
int zero_words(void* p, size_t bytes, hipStream_t st);
const char* const* grammar_names();

// Lexer on device memory.  On success returns 1/0 (valid / rejected) and
// fills count; outputs live in c.lx_*.
int lex_on_device(Context& c, int mode, const uint8_t* d_in, uint64_t n, hipStream_t st, uint64_t* count);

}  // namespace alpacc
