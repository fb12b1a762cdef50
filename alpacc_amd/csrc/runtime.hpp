// Host-side context of one resident grammar (shared by the runtime sources).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <utility>
#include <vector>

#include "grammar.hpp"
#include "lexer.hpp"
#include "timer.hpp"

namespace alpacc {

struct ParseResult;

// Grow-only device allocation: steady-state calls allocate nothing.  gen
// counts the allocations (a captured launch sequence holds the pointer).
// A block outgrown is retired, not freed (hipFree waits for the whole
// device, stalling every other stream's work, such as a server's other
// lanes); retired blocks go with release(), or when an allocation would
// not fit otherwise.  Growth at least doubles, so they add at most the
// current size.
struct DevBuf {
  void* ptr = nullptr;
  size_t cap = 0;
  uint64_t gen = 0;
  std::vector<void*> retired;
  int reserve(size_t bytes);
  void release();
  template <typename T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

// Grow-only pinned host allocation (hipHostMalloc), outgrown blocks retired
// as DevBuf's.
struct PinnedBuf {
  void* ptr = nullptr;
  size_t cap = 0;
  uint64_t gen = 0;
  std::vector<void*> retired;
  int reserve(size_t bytes);
  void release();
  template <typename T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

// Byte layout of a batch's metadata, the same on the device and in the
// pinned host mirror: uploaded [seg_off u64 | chunk_first u32 | tiles u32 |
// inputs that are not padding u32], read back [seg_tok u64 | seg_flags
// u32], then (device only) tile_seg u32[2 * ntiles].
struct BatchLayout {
  size_t off_seg, off_cf, off_nt, off_live, off_tok, off_flags, off_tseg, up_bytes, down_bytes, total;
  BatchLayout(uint64_t count, uint64_t ntiles) {
    off_seg = 0;
    off_cf = (count + 1) * 8;
    off_nt = off_cf + (count + 1) * 4;
    off_live = off_nt + 4;
    up_bytes = off_live + 4;
    off_tok = (up_bytes + 7) & ~size_t(7);
    off_flags = off_tok + (count + 1) * 8;
    down_bytes = off_flags + count * 4 - off_tok;
    off_tseg = (off_flags + count * 4 + 255) & ~size_t(255);
    total = off_tseg + ntiles * 8 + 64;
  }
};

// The capacity a launch sequence is sized for, so that it can be captured
// once in a graph and replayed for every batch within it: `count` inputs
// of `bytes` bytes in total, the first `live` of them real and the rest
// empty padding.  Grids and workspaces follow the capacity, the kernels the
// actual sizes (in device memory); nothing is read back before the
// sequence ends.
struct BatchCaps {
  uint64_t count, bytes, live;
  // tokens the parser is sized for (0: at most bytes); a batch with more
  // flags the launch's overflow word and runs again eagerly
  uint64_t tokens = 0;
};

struct Context {
  const GrammarDesc* g = nullptr;  // the grammar, or gv for a width variant
  GrammarDesc gv{};
  int device = 0;
  hipStream_t stream = nullptr;
  // lexer shape: input bytes per lane (the chunk), predecessor tiles per
  // lookback probe, workgroups per CU of the persistent kernels (0 = as
  // many as fit); lex_grid* are the persistent grids of that shape (0 =
  // not computed yet)
  int lex_ipt = LEX_IPT, lex_lb = LEX_LB_WIN, lex_wpc = 0;
  int lex_grid = 0, lex_grid2 = 0;
  int parse_pt = 1024, parse_apl = 4;  // parser tile shape: positions per tile and per resolve lane
  // seg_resolve's staging (nodes, terminal productions per tile) in a
  // sequence sized for a capacity, learned from its eager run (0: the
  // grammar's maxima)
  int cap_stage_n = 0, cap_stage_t = 0;
  // tokens such a sequence's parser is sized for, learned the same way (0:
  // as many as bytes)
  uint64_t cap_tokens = 0;
  int resolve_br8 = -1;              // 1: every bracket symbol below 256 (-1: not checked yet)
  int resolve_bw = 0;                // bracket stack bytes an entry (0: 1 when resolve_br8, else 2)
  // resident tables
  uint8_t* d_class = nullptr;
  uint16_t* d_step = nullptr;
  uint16_t* d_compose = nullptr;
  uint8_t* d_endo_init = nullptr;
  uint16_t* d_delta_lex = nullptr;
  uint16_t* d_delta_both = nullptr;
  uint8_t* d_token_lex = nullptr;
  uint8_t* d_token_both = nullptr;
  // per window code: pair id | record head << 16 (the parser's pass A reads
  // one entry per position)
  uint64_t* d_wrec = nullptr;
  uint64_t* d_wkeys = nullptr;  // hashed window tables: the code of each slot
  uint8_t* d_pair_rec = nullptr;
  uint16_t* d_prod16 = nullptr;  // two-byte production ids of the pair records (else null)
  // workspaces
  DevBuf small;                       // counters, flags, tickets
  DevBuf lx_term, lx_start, lx_len;   // lexer outputs
  DevBuf lx_status;                   // lookback words of the lexer
  DevBuf in_copy;                     // staging of host inputs
  DevBuf ps_a, ps_b, ps_c, ps_d, ps_e, ps_f, ps_g, ps_h, ps_i;  // parser temporaries
  DevBuf tr_prod, tr_parent, tr_tokpar;                         // tree outputs
  DevBuf lx_meta;     // batch metadata of the lexer (BatchLayout)
  PinnedBuf h_meta;   // its host mirror
  // per-input results of the last lexer launch, in h_meta / lx_meta
  uint64_t lb_count = 0;
  const uint32_t* lb_flags = nullptr;  // [count] LEX_REJECT_* bits
  const uint64_t* lb_tok = nullptr;    // [count + 1] first token of each input; [count] = all
  const uint32_t* d_lb_flags = nullptr;
  const uint64_t* d_lb_tok = nullptr;
  const uint32_t* d_lb_live = nullptr;  // [1] inputs that are not padding
  // parser batches
  DevBuf pmeta;       // offsets of single-string parses
  PinnedBuf h_pmeta;  // tile_first of the parser tiles
  PinnedBuf h_pres;   // per-input results read back
  PinnedBuf h_out;    // batch outputs read back for the wire records
  uint64_t pb_count = 0;
  const struct ParseResult* pb_res = nullptr;
  // host arrays of the last device batch (alpacc_hip_lex_parse_batch_device)
  std::vector<uint8_t> bh_valid;
  std::vector<uint64_t> bh_node_off, bh_nodes, bh_tp_off;
  uint64_t* host_small = nullptr;    // pinned: count / flags read-back
  // optional per-kernel timing with events on the launch stream
  bool timing = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double lex_ms = 0, parse_ms = 0;
  uint64_t lex_launches = 0, parse_launches = 0;
  KernelTimer timer;

  // a lexer stream (alpacc_hip_lex_stream_*): the endomorphism of its bytes
  // so far, their count, where its open token started, its LEX_REJECT_* bits
  struct LexStream {
    bool on = false;
    int mode = 0;
    uint32_t endo = 0;
    uint64_t pos = 0;
    int64_t open_start = 0;
    uint32_t flags = 0;
  } lstream;
  DevBuf lx_stream;  // [4] words a stream's launch reports
  // the combined lex -> parse -> tree sequence captured for a batch
  // capacity (graph_caps), replayed for batches within it that read the
  // same inputs on the same stream while its buffers stay where they were
  bool use_graphs = false;
  // while capturing, the batch functions enqueue device work only and note
  // here the copies between host and device that the graph's runner makes
  // around each launch
  struct Copy {
    void* dst = nullptr;
    const void* src = nullptr;
    size_t bytes = 0;
    hipMemcpyKind kind = hipMemcpyHostToDevice;
  };
  bool capturing = false;
  Copy cp_meta_up, cp_lex_down, cp_parse_down;
  // what a replay restores (calls between replays reuse the fields)
  struct GraphState {
    Copy up, lex_down, parse_down;
    uint64_t lb_count = 0, pb_count = 0;
    const uint32_t* lb_flags = nullptr;
    const uint64_t* lb_tok = nullptr;
    const uint32_t* d_lb_flags = nullptr;
    const uint64_t* d_lb_tok = nullptr;
    const uint32_t* d_lb_live = nullptr;
    const struct ParseResult* pb_res = nullptr;
  } graph_state;
  hipGraphExec_t graph = nullptr;
  BatchCaps graph_caps{};
  const uint8_t* graph_in = nullptr;
  hipStream_t graph_stream = nullptr;
  uint64_t graph_gen = 0;
  uint64_t graph_captures = 0, graph_replays = 0;

  std::vector<DevBuf*> all_buffers() {
    return {&small, &lx_term, &lx_start, &lx_len,    &lx_status, &in_copy, &ps_a,      &ps_b,     &ps_c,  &ps_d, &ps_e,
            &ps_f,  &ps_g,    &ps_h,     &ps_i,      &tr_prod,   &tr_parent, &tr_tokpar, &lx_meta, &pmeta, &lx_stream};
  }
  // allocations of the buffers that the captured sequence uses
  uint64_t graph_buffers_gen() const {
    return small.gen + lx_term.gen + lx_start.gen + lx_len.gen + lx_status.gen + ps_a.gen + ps_b.gen + ps_c.gen +
           tr_prod.gen + tr_parent.gen + tr_tokpar.gen + lx_meta.gen + h_meta.gen + h_pmeta.gen + h_pres.gen;
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
// Zero-fills up to ZERO_RANGES_MAX ranges of (pointer, bytes; 4-byte words)
// in one launch.
constexpr int ZERO_RANGES_MAX = 4;
int zero_ranges(std::initializer_list<std::pair<void*, size_t>> ranges, hipStream_t st);
// Raises kernel k's dynamic-LDS limit on the current device to at least
// `bytes`.  The limit belongs to the kernel for the whole process, so the
// grants of every context are kept in one mutex-guarded table per (device,
// kernel) and only ever raised: a context of a smaller grammar never lowers
// what another context's launches rely on.
hipError_t grant_lds(const void* k, size_t bytes);

// Lexer over `count` inputs packed in device memory at d_in, input i being
// bytes [h_off[i], h_off[i + 1]) (host array).  Returns 0 or an error;
// per-input verdicts and token offsets are in c.lb_* (host) and c.d_lb_*
// (device), the tokens in c.lx_* with starts relative to their input.  With
// caps, the host copies are written when the stream reaches them.  Without
// `enqueue` only the host metadata of the batch is written (the copy that
// a captured sequence uploads).
// One launch of a stream (alpacc_hip_lex_stream_push): the fields of
// LexParams that carry the stream's earlier bytes.
struct LexStreamStep {
  uint32_t carry_endo;
  int64_t carry_start, start_bias;
  int is_final;
  uint32_t* out;  // [4] device words
};

int lex_batch_on_device(Context& c, int mode, const uint8_t* d_in, const uint64_t* h_off, uint64_t count,
                        hipStream_t st, const BatchCaps* caps = nullptr, bool enqueue = true,
                        const LexStreamStep* stream = nullptr);

// Wire records of a batch of host payloads, concatenated into one buffer,
// malloc'd or (alloc_fn) taken from alloc_fn(alloc_ctx, bytes); rec_off (if
// not null) receives the start of every input's record.
int batch_records(Context& c, int mode, const uint8_t* packed, const uint64_t* off, uint64_t count, uint8_t** out,
                  size_t* out_len, std::vector<size_t>* rec_off, uint8_t* (*alloc_fn)(void*, size_t) = nullptr,
                  void* alloc_ctx = nullptr);

// One input: returns 1/0 (valid / rejected) and fills count.
int lex_on_device(Context& c, int mode, const uint8_t* d_in, uint64_t n, hipStream_t st, uint64_t* count);

}  // namespace alpacc
