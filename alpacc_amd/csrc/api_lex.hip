// Lexer entry points of the C ABI (lex_string, the device form, and the
// lexer-only wire record).
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/alpacc_hip.h"
#include "lexer.hpp"
#include "runtime.hpp"

namespace alpacc {

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return fail(ALPACC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

constexpr uint64_t LEX_MAX_INPUT = LEX_MAX_LAUNCH_BYTES;

int lex_batch_on_device(Context& c, int mode, const uint8_t* d_in, const uint64_t* h_off, uint64_t count,
                        hipStream_t st, const BatchCaps* caps, bool enqueue, const LexStreamStep* stream) {
  const GrammarDesc* g = c.g;
  if (count == 0 || count > 0xFFFFFFFEull) return fail(ALPACC_ERR_ARG, "batch of %llu inputs", (unsigned long long)count);
  const uint64_t total = h_off[count] - h_off[0];
  if (caps && (count != caps->count || h_off[count] > caps->bytes))
    return fail(ALPACC_ERR_ARG, "batch beyond the capacity its launch sequence is sized for");
  if (h_off[count] > LEX_MAX_INPUT)
    return fail(ALPACC_ERR_LIMIT, "inputs of %llu bytes exceed the %llu-byte launch limit",
                (unsigned long long)h_off[count], (unsigned long long)LEX_MAX_INPUT);
  // lane prefixes keep a flag in bit 15 of their endomorphism ids
  if (g->nendo > 0x8000) return fail(ALPACC_ERR_LIMIT, "%d lexer endomorphisms exceed 32768", g->nendo);
  // every input owns whole chunks of ipt bytes, at least one
  const uint64_t ipt = (uint64_t)c.lex_ipt;
  const int ish = __builtin_ctzll(ipt);  // 64 or 128 bytes a chunk
  uint64_t nchunks = 0;
  for (uint64_t i = 0; i < count; ++i) {
    if (h_off[i + 1] < h_off[i]) return fail(ALPACC_ERR_ARG, "offsets decrease at input %llu", (unsigned long long)i);
    const uint64_t n = h_off[i + 1] - h_off[i];
    nchunks += n ? (n + ipt - 1) >> ish : 1;
  }
  if (nchunks > 0x7FFFFFFFull) return fail(ALPACC_ERR_LIMIT, "too many lexer chunks");
  const uint64_t nt_actual = (nchunks + LEX_WG - 1) / LEX_WG;
  // at capacity: at most bytes / ipt full chunks plus one partial (or
  // empty) chunk per input
  const uint64_t ntiles = caps ? (caps->bytes / ipt + caps->count + LEX_WG - 1) / LEX_WG : nt_actual;
  const BatchLayout L(count, ntiles);
  int r;
  if ((r = c.lx_meta.reserve(L.total)) || (r = c.h_meta.reserve(L.total))) return r;
  uint8_t* hm = c.h_meta.as<uint8_t>();
  uint8_t* dm = c.lx_meta.as<uint8_t>();
  uint64_t* h_seg = reinterpret_cast<uint64_t*>(hm + L.off_seg);
  uint32_t* h_cf = reinterpret_cast<uint32_t*>(hm + L.off_cf);
  uint32_t k = 0;
  for (uint64_t i = 0; i < count; ++i) {
    h_seg[i] = h_off[i];
    h_cf[i] = k;
    const uint64_t n = h_off[i + 1] - h_off[i];
    k += n ? (uint32_t)((n + ipt - 1) >> ish) : 1u;
  }
  h_seg[count] = h_off[count];
  h_cf[count] = k;
  *reinterpret_cast<uint32_t*>(hm + L.off_nt) = (uint32_t)nt_actual;
  *reinterpret_cast<uint32_t*>(hm + L.off_live) = (uint32_t)(caps ? caps->live : count);
  if (!enqueue) return 0;

  const int lb = g->length_bytes;
  const uint64_t cap = (caps ? caps->bytes : total) + 1;
  // scan state: count_status u64[ntiles] | tile_endo u32[ntiles] | tile_prefix u16[ntiles] |
  // tile_blk u32[2 * scan blocks] | lane_prefix
  const size_t nblk = ((size_t)ntiles + 255) / 256;
  const size_t st_bytes = (size_t)ntiles * 8, te_off = st_bytes, tp_off = te_off + (size_t)ntiles * 4,
               tb_off = (tp_off + (size_t)ntiles * 2 + 15) & ~size_t(15),
               lp_off = (tb_off + nblk * 8 + 255) & ~size_t(255);
  if ((r = c.lx_term.reserve(cap)) || (r = c.lx_start.reserve(cap * 8)) || (r = c.lx_len.reserve(cap * lb)) ||
      (r = c.lx_status.reserve(lp_off + (size_t)ntiles * LEX_WG * 2 + 64)))
    return r;
  LexParams p{};
  p.in = d_in;
  p.in_bytes = h_off[count];
  p.seg_off = reinterpret_cast<const uint64_t*>(dm + L.off_seg);
  p.chunk_first = reinterpret_cast<const uint32_t*>(dm + L.off_cf);
  p.nt_dev = reinterpret_cast<const uint32_t*>(dm + L.off_nt);
  p.tile_seg = reinterpret_cast<uint32_t*>(dm + L.off_tseg);
  p.nseg = (uint32_t)count;
  p.byte_class = c.d_class;
  p.step = c.d_step;
  p.compose = c.d_compose;
  p.endo_init = c.d_endo_init;
  const bool lex_ids = mode == ALPACC_MODE_LEXER;
  p.delta = lex_ids ? c.d_delta_lex : c.d_delta_both;
  p.token = lex_ids ? c.d_token_lex : c.d_token_both;
  p.nstate = g->nstate;
  p.nendo = g->nendo;
  p.nclass = g->nclass;
  p.ignore = lex_ids ? g->ignore_lex : g->ignore_both;
  p.dead = lex_ids ? g->dead_lex : g->dead_both;
  p.max_len = lb >= 8 ? ~0ull : ((1ull << (8 * lb)) - 1);
  p.length_bytes = lb;
  p.terminals = c.lx_term.as<uint8_t>();
  p.starts = c.lx_start.as<int64_t>();
  p.lengths = c.lx_len.ptr;
  p.seg_flags = reinterpret_cast<uint32_t*>(dm + L.off_flags);
  p.seg_tok = reinterpret_cast<uint64_t*>(dm + L.off_tok);
  p.ticket = reinterpret_cast<uint32_t*>(c.small.as<uint64_t>() + 1) + 1;
  uint8_t* sb = c.lx_status.as<uint8_t>();
  p.count_status = reinterpret_cast<uint64_t*>(sb);
  p.tile_endo = reinterpret_cast<uint32_t*>(sb + te_off);
  p.tile_prefix = reinterpret_cast<uint16_t*>(sb + tp_off);
  p.tile_blk = reinterpret_cast<uint32_t*>(sb + tb_off);
  p.lane_prefix = reinterpret_cast<uint16_t*>(sb + lp_off);
  p.ntiles = (uint32_t)ntiles;
  p.ipt = c.lex_ipt;
  p.lb_win = c.lex_lb;
  p.is_final = 1;
  if (stream) {
    p.carry_endo = stream->carry_endo;
    p.carry_start = stream->carry_start;
    p.start_bias = stream->start_bias;
    p.is_final = stream->is_final;
    p.stream_out = stream->out;
  }
  if (!c.lex_grid) HIP_TRY(lex_grid_size(p, c.lex_wpc, &c.lex_grid, &c.lex_grid2));
  c.cp_meta_up = Context::Copy{dm, hm, L.up_bytes, hipMemcpyHostToDevice};
  if (!c.capturing) HIP_TRY(hipMemcpyAsync(dm, hm, L.up_bytes, hipMemcpyHostToDevice, st));
  if ((r = zero_ranges({{c.small.ptr, 16}, {c.lx_status.ptr, st_bytes}, {dm + L.off_flags, count * 4}}, st))) return r;
  const bool timed = c.timing && !caps;
  if (timed) HIP_TRY(hipEventRecord(c.ev[0], st));
  HIP_TRY(lex_launch(p, c.lex_grid, c.lex_grid2, st, c.timer));
  if (timed) HIP_TRY(hipEventRecord(c.ev[1], st));
  c.cp_lex_down = Context::Copy{hm + L.off_tok, dm + L.off_tok, L.down_bytes, hipMemcpyDeviceToHost};
  if (!c.capturing) HIP_TRY(hipMemcpyAsync(hm + L.off_tok, dm + L.off_tok, L.down_bytes, hipMemcpyDeviceToHost, st));
  if (!caps) {
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(c.timer.collect());
  }
  if (timed) {
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, c.ev[0], c.ev[1]));
    c.lex_ms += ms;
    c.lex_launches++;
  }
  c.lb_count = count;
  c.lb_tok = reinterpret_cast<const uint64_t*>(hm + L.off_tok);
  c.lb_flags = reinterpret_cast<const uint32_t*>(hm + L.off_flags);
  c.d_lb_tok = reinterpret_cast<const uint64_t*>(dm + L.off_tok);
  c.d_lb_flags = reinterpret_cast<const uint32_t*>(dm + L.off_flags);
  c.d_lb_live = reinterpret_cast<const uint32_t*>(dm + L.off_live);
  return 0;
}

int lex_on_device(Context& c, int mode, const uint8_t* d_in, uint64_t n, hipStream_t st, uint64_t* count) {
  *count = 0;
  if (n == 0) return ALPACC_VALID;
  const uint64_t off[2] = {0, n};
  int r = lex_batch_on_device(c, mode, d_in, off, 1, st);
  if (r < 0) return r;
  *count = c.lb_tok[1] - c.lb_tok[0];
  return c.lb_flags[0] ? ALPACC_REJECT : ALPACC_VALID;
}

// Stage a host buffer on the device (grow-only staging buffer).
int stage_input(Context& c, const void* host, uint64_t bytes, hipStream_t st, const uint8_t** d) {
  int r;
  if ((r = c.in_copy.reserve(bytes + 16))) return r;
  if (bytes) HIP_TRY(hipMemcpyAsync(c.in_copy.ptr, host, bytes, hipMemcpyHostToDevice, st));
  *d = c.in_copy.as<uint8_t>();
  return 0;
}

int lexemes_to_host(Context& c, int mode, uint64_t count, hipStream_t st, alpacc_hip_lexemes* out) {
  const GrammarDesc* g = c.g;
  const int tb = mode == ALPACC_MODE_LEXER ? g->terminal_bytes_lex : g->terminal_bytes_both;
  const int lb = g->length_bytes;
  out->count = count;
  out->terminals = malloc(count * tb + 1);
  out->starts = static_cast<int64_t*>(malloc(count * 8 + 1));
  out->lengths = malloc(count * lb + 1);
  if (!out->terminals || !out->starts || !out->lengths) {
    free(out->terminals);
    free(out->starts);
    free(out->lengths);
    out->terminals = out->starts = nullptr;
    out->lengths = nullptr;
    return fail(ALPACC_ERR_NOMEM, "host allocation failed");
  }
  if (count) {
    HIP_TRY(hipMemcpyAsync(out->terminals, c.lx_term.ptr, count * tb, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(out->starts, c.lx_start.ptr, count * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(out->lengths, c.lx_len.ptr, count * lb, hipMemcpyDeviceToHost, st));
  }
  HIP_TRY(hipStreamSynchronize(st));
  return 0;
}

void lex_stream_begin(Context& c, int mode) {
  c.lstream = Context::LexStream{};
  c.lstream.on = true;
  c.lstream.mode = mode;
}

// One chunk of the stream begun on c (alpacc_hip_lex_stream_push).
int lex_stream_push(Context& c, const uint8_t* d_chunk, uint64_t n, int is_last_chunk, alpacc_hip_lexemes* d_out,
                    hipStream_t st) {
  Context::LexStream& ls = c.lstream;
  const GrammarDesc* g = c.g;
  const bool lex_ids = ls.mode == ALPACC_MODE_LEXER;
  const int lb = g->length_bytes;
  const uint64_t max_len = lb >= 8 ? ~0ull : ((1ull << (8 * lb)) - 1);
  d_out->count = 0;
  d_out->terminals = c.lx_term.ptr;
  d_out->starts = c.lx_start.as<int64_t>();
  d_out->lengths = c.lx_len.ptr;
  int r;
  if (n) {
    // the chunk as one input that continues the stream: its lanes start
    // from the stream's state, its first token from the open one
    if ((r = c.lx_stream.reserve(16))) return r;
    uint32_t* words = c.lx_stream.as<uint32_t>();
    const LexStreamStep step{ls.endo, ls.open_start - (int64_t)ls.pos, (int64_t)ls.pos, is_last_chunk ? 1 : 0, words};
    const uint64_t off[2] = {0, n};
    if ((r = lex_batch_on_device(c, ls.mode, d_chunk, off, 1, st, nullptr, true, &step)) < 0) return r;
    uint32_t* hw = reinterpret_cast<uint32_t*>(c.host_small + 12);  // pinned
    HIP_TRY(hipMemcpyAsync(hw, words, 12, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    ls.flags |= c.lb_flags[0];
    ls.endo = ls.endo ? g->compose[(size_t)ls.endo * g->nendo + hw[0]] : hw[0];
    if (hw[2]) ls.open_start = (int64_t)ls.pos + (int64_t)hw[2] - 1;
    ls.pos += n;
    d_out->count = c.lb_tok[1] - c.lb_tok[0];
    d_out->terminals = c.lx_term.ptr;
    d_out->starts = c.lx_start.as<int64_t>();
    d_out->lengths = c.lx_len.ptr;
  } else if (is_last_chunk && ls.pos) {
    // the stream ended with the previous chunk: its open token and the
    // final-state check, from the state its bytes lead to
    const int state = g->endo_init[ls.endo];
    const int t = (lex_ids ? g->token_lex : g->token_both)[state];
    const int dead = lex_ids ? g->dead_lex : g->dead_both, ignore = lex_ids ? g->ignore_lex : g->ignore_both;
    if (t == dead) {
      ls.flags |= LEX_REJECT_FINAL;
    } else if (t != ignore) {
      const uint64_t len = ls.pos - (uint64_t)ls.open_start;
      if (len > max_len) ls.flags |= LEX_REJECT_OVERFLOW;
      if ((r = c.lx_term.reserve(16)) || (r = c.lx_start.reserve(16)) || (r = c.lx_len.reserve(16))) return r;
      uint8_t* hb = reinterpret_cast<uint8_t*>(c.host_small + 12);  // pinned
      const uint8_t term = (uint8_t)t;
      const int64_t start = ls.open_start;
      memcpy(hb, &term, 1);
      memcpy(hb + 8, &start, 8);
      memcpy(hb + 16, &len, (size_t)lb);  // little-endian: the low lb bytes
      HIP_TRY(hipMemcpyAsync(c.lx_term.ptr, hb, 1, hipMemcpyHostToDevice, st));
      HIP_TRY(hipMemcpyAsync(c.lx_start.ptr, hb + 8, 8, hipMemcpyHostToDevice, st));
      HIP_TRY(hipMemcpyAsync(c.lx_len.ptr, hb + 16, (size_t)lb, hipMemcpyHostToDevice, st));
      HIP_TRY(hipStreamSynchronize(st));
      d_out->count = 1;
      d_out->terminals = c.lx_term.ptr;
      d_out->starts = c.lx_start.as<int64_t>();
      d_out->lengths = c.lx_len.ptr;
    }
  }
  if (is_last_chunk) {
    ls.on = false;
    return ls.flags ? ALPACC_REJECT : ALPACC_VALID;
  }
  return (ls.flags & LEX_REJECT_OVERFLOW) ? ALPACC_REJECT : ALPACC_VALID;
}


int lexer_record_stream(Context& c, uint64_t n, const uint8_t* payload, std::vector<uint8_t>& rec);

// Lexer-only response record: u8 1, u64 count, then (terminal, start,
// length) per lexeme with native widths (C/Generator.hs:88-105).  An input
// beyond one launch's limit goes through the stream path in chunks.
int lexer_record(Context& c, uint64_t n, const uint8_t* payload, std::vector<uint8_t>& rec) {
  if (n > LEX_MAX_INPUT) return lexer_record_stream(c, n, payload, rec);
  hipStream_t st = c.stream;
  const uint8_t* d_in = nullptr;
  int r;
  if ((r = stage_input(c, payload, n, st, &d_in))) return r;
  uint64_t count = 0;
  r = lex_on_device(c, ALPACC_MODE_LEXER, d_in, n, st, &count);
  if (r < 0) return r;
  if (r == ALPACC_REJECT) {
    rec.assign(1, 0);
    return ALPACC_REJECT;
  }
  alpacc_hip_lexemes lx{};
  if ((r = lexemes_to_host(c, ALPACC_MODE_LEXER, count, st, &lx))) return r;
  const GrammarDesc* g = c.g;
  const int tb = g->terminal_bytes_lex, ib = g->index_bytes, lb = g->length_bytes;
  const size_t per = tb + ib + lb;
  rec.resize(1 + 8 + count * per);
  rec[0] = 1;
  memcpy(&rec[1], &count, 8);
  uint8_t* o = rec.data() + 9;
  const uint8_t* t = static_cast<const uint8_t*>(lx.terminals);
  const uint8_t* l = static_cast<const uint8_t*>(lx.lengths);
  for (uint64_t i = 0; i < count; ++i) {
    memcpy(o, t + i * tb, tb);
    memcpy(o + tb, reinterpret_cast<const uint8_t*>(lx.starts + i), ib);
    memcpy(o + tb + ib, l + i * lb, lb);
    o += per;
  }
  free(lx.terminals);
  free(lx.starts);
  free(lx.lengths);
  return ALPACC_VALID;
}

}  // namespace alpacc

using namespace alpacc;

extern "C" {

int alpacc_hip_lex_device(alpacc_hip_ctx* h, int mode, const uint8_t* d_str, uint64_t n, alpacc_hip_lexemes* d_out,
                          void* stream) {
  if (!h || !d_out || (n && !d_str)) return fail(ALPACC_ERR_ARG, "null argument");
  if (mode != ALPACC_MODE_LEXER && mode != ALPACC_MODE_BOTH) return fail(ALPACC_ERR_ARG, "bad lexer mode %d", mode);
  Context& c = h->c;
  HIP_TRY(hipSetDevice(c.device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c.stream;
  uint64_t count = 0;
  int r = lex_on_device(c, mode, d_str, n, st, &count);
  if (r < 0) return r;
  d_out->count = count;
  d_out->terminals = c.lx_term.ptr;
  d_out->starts = c.lx_start.as<int64_t>();
  d_out->lengths = c.lx_len.ptr;
  return r;
}

int alpacc_hip_lex_stream_begin(alpacc_hip_ctx* h, int mode) {
  if (!h) return fail(ALPACC_ERR_ARG, "null argument");
  if (mode != ALPACC_MODE_LEXER && mode != ALPACC_MODE_BOTH) return fail(ALPACC_ERR_ARG, "bad lexer mode %d", mode);
  lex_stream_begin(h->c, mode);
  return 0;
}

int alpacc_hip_lex_stream_push(alpacc_hip_ctx* h, const uint8_t* d_chunk, uint64_t n, int is_last_chunk,
                               alpacc_hip_lexemes* d_out, void* stream) {
  if (!h || !d_out || (n && !d_chunk)) return fail(ALPACC_ERR_ARG, "null argument");
  Context& c = h->c;
  if (!c.lstream.on) return fail(ALPACC_ERR_ARG, "no lexer stream begun on this context");
  HIP_TRY(hipSetDevice(c.device));
  return lex_stream_push(c, d_chunk, n, is_last_chunk, d_out, stream ? static_cast<hipStream_t>(stream) : c.stream);
}

int alpacc_hip_set_lex_shape(alpacc_hip_ctx* h, const alpacc_hip_lex_shape* shape) {
  if (!h || !shape) return fail(ALPACC_ERR_ARG, "null argument");
  if (!lex_shape_ok(shape->items_per_thread, shape->lookback_window) || shape->workgroups_per_cu < 0)
    return fail(ALPACC_ERR_ARG, "lexer shape %d bytes per lane, %d workgroups per CU, window %d is not one of the kernels'",
                shape->items_per_thread, shape->workgroups_per_cu, shape->lookback_window);
  Context& c = h->c;
  HIP_TRY(hipSetDevice(c.device));
  c.lex_ipt = shape->items_per_thread;
  c.lex_wpc = shape->workgroups_per_cu;
  c.lex_lb = shape->lookback_window;
  c.lex_grid = c.lex_grid2 = 0;
  if (c.graph) {  // captured with the old kernels
    (void)hipGraphExecDestroy(c.graph);
    c.graph = nullptr;
  }
  return 0;
}

int alpacc_hip_get_lex_shape(const alpacc_hip_ctx* h, alpacc_hip_lex_shape* shape) {
  if (!h || !shape) return fail(ALPACC_ERR_ARG, "null argument");
  shape->items_per_thread = h->c.lex_ipt;
  shape->workgroups_per_cu = h->c.lex_wpc;
  shape->lookback_window = h->c.lex_lb;
  shape->ms = 0;
  return 0;
}

int alpacc_hip_autotune_lexer(alpacc_hip_ctx* h, int mode, const uint8_t* d_sample, uint64_t n, int reps,
                              alpacc_hip_lex_shape* best) {
  if (!h || !best || !d_sample || n == 0 || reps <= 0) return fail(ALPACC_ERR_ARG, "null argument or empty sample");
  if (mode != ALPACC_MODE_LEXER && mode != ALPACC_MODE_BOTH) return fail(ALPACC_ERR_ARG, "bad lexer mode %d", mode);
  Context& c = h->c;
  HIP_TRY(hipSetDevice(c.device));
  for (hipEvent_t& ev : c.ev)
    if (!ev) HIP_TRY(hipEventCreate(&ev));
  const bool timing = c.timing, ktiming = c.timer.on;
  c.timing = true;  // the lexer's stage events time each run
  c.timer.on = false;
  static const int ipts[] = {64, 128}, wpcs[] = {2, 3, 4, 0}, wins[] = {4, 8, 16, 32, 64};
  alpacc_hip_lex_shape top{c.lex_ipt, c.lex_wpc, c.lex_lb, 0};
  double top_ms = -1;
  int r = 0;
  for (int ipt : ipts)
    for (int wpc : wpcs)
      for (int win : wins) {
        const alpacc_hip_lex_shape sh{ipt, wpc, win, 0};
        if ((r = alpacc_hip_set_lex_shape(h, &sh)) < 0) break;
        uint64_t count = 0;
        if ((r = lex_on_device(c, mode, d_sample, n, c.stream, &count)) < 0) break;  // warm
        c.lex_ms = 0;
        c.lex_launches = 0;
        for (int i = 0; i < reps && r >= 0; ++i) r = lex_on_device(c, mode, d_sample, n, c.stream, &count);
        if (r < 0) break;
        const double ms = c.lex_ms / (double)reps;
        if (top_ms < 0 || ms < top_ms) {
          top_ms = ms;
          top = sh;
        }
      }
  c.timing = timing;
  c.timer.on = ktiming;
  c.lex_ms = 0;
  c.lex_launches = 0;
  if (r < 0) return r;
  top.ms = top_ms;
  if ((r = alpacc_hip_set_lex_shape(h, &top)) < 0) return r;
  *best = top;
  return 0;
}

int alpacc_hip_lex_string(alpacc_hip_ctx* h, int mode, const uint8_t* str, uint64_t n, alpacc_hip_lexemes* out) {
  if (!h || !out || (n && !str)) return fail(ALPACC_ERR_ARG, "null argument");
  if (mode != ALPACC_MODE_LEXER && mode != ALPACC_MODE_BOTH) return fail(ALPACC_ERR_ARG, "bad lexer mode %d", mode);
  Context& c = h->c;
  HIP_TRY(hipSetDevice(c.device));
  memset(out, 0, sizeof *out);
  const uint8_t* d_in = nullptr;
  int r;
  if ((r = stage_input(c, str, n, c.stream, &d_in))) return r;
  uint64_t count = 0;
  r = lex_on_device(c, mode, d_in, n, c.stream, &count);
  if (r != ALPACC_VALID) return r;
  if ((r = lexemes_to_host(c, mode, count, c.stream, out))) return r;
  return ALPACC_VALID;
}

}  // extern "C"

namespace alpacc {

// The lexer-only record of an input of any size: its bytes are staged and
// lexed in chunks of LEX_STREAM_CHUNK through the stream entry points, the
// tokens of every chunk appended to the record as they close.

int lexer_record_stream(Context& c, uint64_t n, const uint8_t* payload, std::vector<uint8_t>& rec) {
  hipStream_t st = c.stream;
  const GrammarDesc* g = c.g;
  const int tb = g->terminal_bytes_lex, ib = g->index_bytes, lb = g->length_bytes;
  const size_t per = tb + ib + lb;
  lex_stream_begin(c, ALPACC_MODE_LEXER);
  int r;
  rec.assign(9, 0);
  uint64_t count = 0;
  std::vector<uint8_t> ht, hs, hl;
  for (uint64_t at = 0; at < n;) {
    const uint64_t m = n - at < LEX_STREAM_CHUNK ? n - at : LEX_STREAM_CHUNK;
    const bool last = at + m == n;
    const uint8_t* d_in = nullptr;
    if ((r = stage_input(c, payload + at, m, st, &d_in))) return r;
    alpacc_hip_lexemes d{};
    r = lex_stream_push(c, d_in, m, last ? 1 : 0, &d, st);
    if (r < 0) return r;
    if (r == ALPACC_REJECT) {
      c.lstream.on = false;
      rec.assign(1, 0);
      return ALPACC_REJECT;
    }
    if (d.count) {
      ht.resize(d.count * tb);
      hs.resize(d.count * 8);
      hl.resize(d.count * lb);
      HIP_TRY(hipMemcpyAsync(ht.data(), d.terminals, d.count * tb, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipMemcpyAsync(hs.data(), d.starts, d.count * 8, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipMemcpyAsync(hl.data(), d.lengths, d.count * lb, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      const size_t o0 = rec.size();
      rec.resize(o0 + d.count * per);
      uint8_t* o = rec.data() + o0;
      for (uint64_t i = 0; i < d.count; ++i, o += per) {
        memcpy(o, &ht[i * tb], tb);
        memcpy(o + tb, &hs[i * 8], ib);
        memcpy(o + tb + ib, &hl[i * lb], lb);
      }
      count += d.count;
    }
    at += m;
  }
  rec[0] = 1;
  memcpy(&rec[1], &count, 8);
  return ALPACC_VALID;
}

}  // namespace alpacc
