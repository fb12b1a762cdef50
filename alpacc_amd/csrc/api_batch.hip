// Batch entry points of the C ABI: many inputs, one launch sequence.
//
// The reference answers a counted batch one request at a time (the
// generated run_case per frame, Generator/C/Generator.hs:64-152, driven by
// backends/c/cli.c:167-189).  Here the inputs of a batch are packed back to
// back, every stage runs once over all of them (the lexer's tiles shared by
// the inputs' chunks, the parser's per input), and the records are cut
// from the batch's outputs afterwards: each record is
// byte-identical to what alpacc_hip_run_case returns for that input alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/alpacc_hip.h"
#include "parser.hpp"
#include "runtime.hpp"

namespace alpacc {

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return fail(ALPACC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

int stage_input(Context& c, const void* host, uint64_t bytes, hipStream_t st, const uint8_t** d);
int lexer_record(Context& c, uint64_t n, const uint8_t* payload, std::vector<uint8_t>& rec);
int parser_record(Context& c, uint64_t n, const uint8_t* payload, std::vector<uint8_t>& rec);
int both_record(Context& c, uint64_t n, const uint8_t* payload, std::vector<uint8_t>& rec);

namespace {


// The combined sequence over a batch of `count` inputs at d_in, sized for
// the capacity `caps` (count = caps.count, padded), enqueued on st.
int lex_parse_enqueue_caps(Context& c, const uint8_t* d_in, const uint64_t* off, const BatchCaps& caps, hipStream_t st) {
  int r = lex_batch_on_device(c, ALPACC_MODE_BOTH, d_in, off, caps.count, st, &caps);
  if (r < 0) return r;
  return parse_batch_on_device(c, c.lx_term.as<uint8_t>(), c.d_lb_tok, nullptr, nullptr, caps.count, true, st, &caps,
                               c.d_lb_flags, c.d_lb_live);
}

uint64_t pow2_at_least(uint64_t x, uint64_t lo) {
  uint64_t p = lo;
  while (p < x) p <<= 1;
  return p;
}

// Lexer + parser over a batch in device memory through a graph: the
// sequence is captured once for a capacity (inputs and bytes rounded up to
// powers of two, growing with the batches seen) and replayed for every
// batch within it that reads the same inputs pointer on the same stream,
// while the buffers it uses stay where they were; otherwise it is captured
// again.  A capture first runs the sequence eagerly (sizing every buffer,
// producing this batch's results), then records it.  Fills c.lb_* and
// c.pb_res for the first `count` inputs, like lex_parse_batch.
int lex_parse_batch_graph(Context& c, const uint8_t* d_in, const uint64_t* off, uint64_t count, hipStream_t st) {
  if (count == 0 || count > 0x7FFFFFFFull) return fail(ALPACC_ERR_ARG, "batch of %llu inputs", (unsigned long long)count);
  for (uint64_t i = 0; i < count; ++i)
    if (off[i + 1] < off[i]) return fail(ALPACC_ERR_ARG, "offsets decrease at input %llu", (unsigned long long)i);
  const uint64_t bytes = off[count];
  BatchCaps caps{pow2_at_least(count, 16), pow2_at_least(bytes, 1 << 16), count, c.cap_tokens};
  if (c.graph) {  // a larger capacity already captured serves this batch too
    caps.count = caps.count > c.graph_caps.count ? caps.count : c.graph_caps.count;
    caps.bytes = caps.bytes > c.graph_caps.bytes ? caps.bytes : c.graph_caps.bytes;
  }
  // a capacity beyond the launch limits (lexer bytes, the parser's 31-bit
  // depths at worst-case positions) runs eagerly instead
  const GrammarDesc* g = c.g;
  const uint64_t per_pos = (uint64_t)parse_per_position(g);
  if (caps.bytes > 0x7FFFFFFFull || (caps.bytes + 2 * caps.count) * per_pos + caps.count > 0x7FFFFFFFull) {
    int r = lex_batch_on_device(c, ALPACC_MODE_BOTH, d_in, off, count, st);
    if (r < 0) return r;
    return parse_batch_on_device(c, c.lx_term.as<uint8_t>(), c.d_lb_tok, c.lb_tok, c.lb_flags, count, true, st);
  }
  std::vector<uint64_t> padded(caps.count + 1);
  for (uint64_t i = 0; i <= caps.count; ++i) padded[i] = off[i < count ? i : count];
  const bool timing = c.timer.on;
  c.timer.on = false;  // no event pairs inside a graph
  int r = 0;
  const bool reuse = c.graph && caps.count == c.graph_caps.count && caps.bytes == c.graph_caps.bytes &&
                     caps.tokens == c.graph_caps.tokens &&
                     c.graph_in == d_in && c.graph_stream == st && c.graph_gen == c.graph_buffers_gen();
  auto copy = [&](const Context::Copy& x) {
    return x.bytes ? hipMemcpyAsync(x.dst, x.src, x.bytes, x.kind, st) : hipSuccess;
  };
  Context::GraphState& gs = c.graph_state;
  if (reuse) {
    // the batch's metadata (host block), its upload, the graph, the results
    r = lex_batch_on_device(c, ALPACC_MODE_BOTH, d_in, padded.data(), caps.count, st, &caps, false);
    if (r >= 0) {
      hipError_t e = copy(gs.up);
      if (e == hipSuccess) e = hipGraphLaunch(c.graph, st);
      if (e == hipSuccess) e = copy(gs.lex_down);
      // the scan's totals and the results of the live inputs, in one copy
      Context::Copy pd = gs.parse_down;
      pd.bytes -= (caps.count - count) * sizeof(ParseResult);
      if (e == hipSuccess) e = copy(pd);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) r = fail(ALPACC_ERR_HIP, "graph replay: %s", hipGetErrorString(e));
      c.graph_replays++;
      // the results where the captured sequence leaves them
      c.lb_count = gs.lb_count;
      c.lb_flags = gs.lb_flags;
      c.lb_tok = gs.lb_tok;
      c.d_lb_flags = gs.d_lb_flags;
      c.d_lb_tok = gs.d_lb_tok;
      c.d_lb_live = gs.d_lb_live;
      c.pb_count = gs.pb_count;
      c.pb_res = gs.pb_res;
      if (r >= 0 && parse_scan_overflow(c)) {
        // a tile beyond the staging learned from earlier batches: this
        // batch runs again eagerly, and the next one captures again with
        // staging for the grammar's maxima
        (void)hipGraphExecDestroy(c.graph);
        c.graph = nullptr;
        c.cap_stage_n = c.cap_stage_t = 0;
        c.cap_tokens = 0;
        c.timer.on = timing;
        r = lex_batch_on_device(c, ALPACC_MODE_BOTH, d_in, off, count, st);
        if (r >= 0) r = parse_batch_on_device(c, c.lx_term.as<uint8_t>(), c.d_lb_tok, c.lb_tok, c.lb_flags, count, true, st);
        return r < 0 ? r : 0;
      }
    }
  } else {
    if (c.graph) {
      (void)hipGraphExecDestroy(c.graph);
      c.graph = nullptr;
    }
    // eager run: sizes every buffer and answers this batch (with staging
    // for the grammar's maxima when the learned one falls short)
    auto eager = [&]() {
      int rr = lex_parse_enqueue_caps(c, d_in, padded.data(), caps, st);
      if (rr >= 0) {
        const hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess) rr = fail(ALPACC_ERR_HIP, "batch: %s", hipGetErrorString(e));
      }
      return rr;
    };
    // the eager run sizes the parser for the bytes (as many tokens as
    // bytes), so every buffer the captured sequence uses, with the token
    // capacity learned below, is already there when the capture starts
    caps.tokens = 0;
    r = eager();
    if (r >= 0 && parse_scan_overflow(c)) {
      c.cap_stage_n = c.cap_stage_t = 0;
      r = eager();
    }
    if (r >= 0) {
      // the captured sequence stages a quarter more than this batch's
      // largest tile needs (a later batch beyond it runs again eagerly)
      int n = 0, t = 0;
      parse_scan_maxima(c, true, &n, &t);
      c.cap_stage_n = (n + n / 4 + 64) & ~63;
      c.cap_stage_t = (t + t / 4 + 64) & ~63;
      // and the parser is sized for half again this batch's tokens, to a
      // power of two (the grids and workspaces of a bound by bytes are
      // several times what a batch of JSON needs)
      const uint64_t toks = c.lb_tok[count] - c.lb_tok[0];
      c.cap_tokens = pow2_at_least(toks + toks / 2, 1 << 16);
      caps.tokens = c.cap_tokens < caps.bytes ? c.cap_tokens : 0;
      c.cap_tokens = caps.tokens;
    }
    if (r >= 0) {
      const uint64_t gen = c.graph_buffers_gen();
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        c.capturing = true;  // device work only: the copies run around each launch
        r = lex_parse_enqueue_caps(c, d_in, padded.data(), caps, st);
        c.capturing = false;
        e = hipStreamEndCapture(st, &graph);
      }
      if (r >= 0 && e == hipSuccess && c.graph_buffers_gen() != gen) r = fail(ALPACC_ERR_HIP, "buffers moved during capture");
      if (r >= 0 && e == hipSuccess) e = hipGraphInstantiate(&c.graph, graph, nullptr, nullptr, 0);
      if (graph) (void)hipGraphDestroy(graph);
      if (r >= 0 && e != hipSuccess) r = fail(ALPACC_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
      if (r >= 0) {
        c.graph_caps = caps;
        c.graph_in = d_in;
        c.graph_stream = st;
        c.graph_gen = gen;
        c.graph_captures++;
        gs.up = c.cp_meta_up;
        gs.lex_down = c.cp_lex_down;
        gs.parse_down = c.cp_parse_down;
        gs.lb_count = c.lb_count;
        gs.lb_flags = c.lb_flags;
        gs.lb_tok = c.lb_tok;
        gs.d_lb_flags = c.d_lb_flags;
        gs.d_lb_tok = c.d_lb_tok;
        gs.d_lb_live = c.d_lb_live;
        gs.pb_count = c.pb_count;
        gs.pb_res = c.pb_res;
      } else if (c.graph) {
        (void)hipGraphExecDestroy(c.graph);
        c.graph = nullptr;
      }
    }
  }
  c.timer.on = timing;
  return r < 0 ? r : 0;
}

// Lexer + parser over a batch in device memory: fills c.lb_* and c.pb_res
// (through the captured graph when the context has graphs on).
int lex_parse_batch(Context& c, const uint8_t* d_in, const uint64_t* off, uint64_t count, hipStream_t st) {
  if (c.use_graphs) return lex_parse_batch_graph(c, d_in, off, count, st);
  int r = lex_batch_on_device(c, ALPACC_MODE_BOTH, d_in, off, count, st);
  if (r < 0) return r;
  return parse_batch_on_device(c, c.lx_term.as<uint8_t>(), c.d_lb_tok, c.lb_tok, c.lb_flags, count, true, st);
}

// Device arrays copied into one pinned host buffer (c.h_out), in order.
struct Pull {
  const void* d;
  size_t bytes;
  const uint8_t* h = nullptr;
};

int pull_all(Context& c, Pull* v, int n, hipStream_t st) {
  size_t tot = 0;
  for (int i = 0; i < n; ++i) tot += (v[i].bytes + 255) & ~size_t(255);
  int r;
  if ((r = c.h_out.reserve(tot + 256))) return r;
  uint8_t* h = c.h_out.as<uint8_t>();
  for (int i = 0; i < n; ++i) {
    if (v[i].bytes) HIP_TRY(hipMemcpyAsync(h, v[i].d, v[i].bytes, hipMemcpyDeviceToHost, st));
    v[i].h = h;
    h += (v[i].bytes + 255) & ~size_t(255);
  }
  HIP_TRY(hipStreamSynchronize(st));
  return 0;
}

}  // namespace

// Records of every input, concatenated into one malloc'd buffer (*out,
// *out_len); rec_off (if not null) receives the start of every record.
// `off` is rebased so that input i is bytes [off[i] - off[0],
// off[i + 1] - off[0]) of `packed`.
int batch_records(Context& c, int mode, const uint8_t* packed, const uint64_t* off_in, uint64_t count, uint8_t** out,
                  size_t* out_len, std::vector<size_t>* rec_off, uint8_t* (*alloc_fn)(void*, size_t), void* alloc_ctx) {
  hipStream_t st = c.stream;
  const GrammarDesc* g = c.g;
  *out = nullptr;
  *out_len = 0;
  std::vector<uint64_t> off(count + 1);
  for (uint64_t i = 0; i <= count; ++i) off[i] = off_in[i] - off_in[0];
  const uint8_t* d_in = nullptr;
  int r;
  if ((r = stage_input(c, packed + off_in[0], off[count], st, &d_in))) return r;
  if (rec_off) rec_off->resize(count);
  const int ib = g->index_bytes, lb = g->length_bytes;
  auto alloc = [&](size_t n) -> uint8_t* {
    uint8_t* p = alloc_fn ? alloc_fn(alloc_ctx, n ? n : 1) : static_cast<uint8_t*>(malloc(n ? n : 1));
    if (p) {
      *out = p;
      *out_len = n;
    }
    return p;
  };
  if (mode == ALPACC_MODE_LEXER) {
    if ((r = lex_batch_on_device(c, mode, d_in, off.data(), count, st)) < 0) return r;
    const uint64_t T = c.lb_tok[count];
    const int tb = g->terminal_bytes_lex;
    Pull v[3] = {{c.lx_term.ptr, T * tb}, {c.lx_start.ptr, T * 8}, {c.lx_len.ptr, T * lb}};
    if ((r = pull_all(c, v, 3, st))) return r;
    size_t total = 0;
    for (uint64_t i = 0; i < count; ++i)
      total += c.lb_flags[i] ? 1 : 9 + (c.lb_tok[i + 1] - c.lb_tok[i]) * (tb + ib + lb);
    uint8_t* o = alloc(total);
    if (!o) return fail(ALPACC_ERR_NOMEM, "host allocation failed");
    for (uint64_t i = 0; i < count; ++i) {
      if (rec_off) (*rec_off)[i] = o - *out;
      if (c.lb_flags[i]) {
        *o++ = 0;
        continue;
      }
      const uint64_t a = c.lb_tok[i], n = c.lb_tok[i + 1] - a;
      *o++ = 1;
      memcpy(o, &n, 8);
      o += 8;
      for (uint64_t k = a; k < a + n; ++k) {
        memcpy(o, v[0].h + k * tb, tb);
        memcpy(o + tb, v[1].h + k * 8, ib);
        memcpy(o + tb + ib, v[2].h + k * lb, lb);
        o += tb + ib + lb;
      }
    }
    return 0;
  }
  if (mode == ALPACC_MODE_PARSER) {
    if (g->terminal_bytes_both != 1) return fail(ALPACC_ERR_LIMIT, "parser batches take u8 terminal ids");
    if ((r = c.pmeta.reserve((count + 1) * 8 + 64)) || (r = c.h_meta.reserve((count + 1) * 8 + 64))) return r;
    uint64_t* h_off = c.h_meta.as<uint64_t>();
    memcpy(h_off, off.data(), (count + 1) * 8);
    HIP_TRY(hipMemcpyAsync(c.pmeta.ptr, h_off, (count + 1) * 8, hipMemcpyHostToDevice, st));
    if ((r = parse_batch_on_device(c, d_in, c.pmeta.as<uint64_t>(), h_off, nullptr, count, false, st)) < 0)
      return r;
    const uint64_t pb = (uint64_t)c.g->production_bytes;
    uint64_t np = 0;
    size_t total = 0;
    for (uint64_t i = 0; i < count; ++i) {
      const ParseResult& o = c.pb_res[i];
      if (!o.flags && o.slot_off + o.slots > np) np = o.slot_off + o.slots;
      total += o.flags ? 1 : 9 + o.slots * pb;
    }
    Pull v[1] = {{c.ps_d.ptr, np * pb}};
    if ((r = pull_all(c, v, 1, st))) return r;
    uint8_t* w = alloc(total);
    if (!w) return fail(ALPACC_ERR_NOMEM, "host allocation failed");
    for (uint64_t i = 0; i < count; ++i) {
      const ParseResult& o = c.pb_res[i];
      if (rec_off) (*rec_off)[i] = w - *out;
      if (o.flags) {
        *w++ = 0;
        continue;
      }
      *w++ = 1;
      memcpy(w, &o.slots, 8);
      w += 8;
      memcpy(w, v[0].h + o.slot_off * pb, o.slots * pb);
      w += o.slots * pb;
    }
    return 0;
  }
  // combined
  if ((r = lex_parse_batch(c, d_in, off.data(), count, st)) < 0) return r;
  const uint64_t T = c.lb_tok[count];
  const int tb = g->terminal_bytes_both, pb = g->production_bytes;
  uint64_t nodes = 0, terms = 0;
  size_t total = 0;
  for (uint64_t i = 0; i < count; ++i) {
    const ParseResult& o = c.pb_res[i];
    if (c.lb_flags[i] || o.flags) {
      total += 1;
      continue;
    }
    const uint64_t n = c.lb_tok[i + 1] - c.lb_tok[i];
    nodes = o.node_off + o.nodes > nodes ? o.node_off + o.nodes : nodes;
    terms = o.term_off + n > terms ? o.term_off + n : terms;
    total += 1 + 8 + n * (tb + ib + lb) + 8 + o.nodes * (pb + ib) + n * ib;
  }
  Pull v[6] = {{c.lx_term.ptr, T * tb},       {c.lx_start.ptr, T * 8},     {c.lx_len.ptr, T * lb},
               {c.tr_prod.ptr, nodes * pb},   {c.tr_parent.ptr, nodes * 8}, {c.tr_tokpar.ptr, terms * 8}};
  if ((r = pull_all(c, v, 6, st))) return r;
  uint8_t* w = alloc(total);
  if (!w) return fail(ALPACC_ERR_NOMEM, "host allocation failed");
  for (uint64_t i = 0; i < count; ++i) {
    const ParseResult& o = c.pb_res[i];
    if (rec_off) (*rec_off)[i] = w - *out;
    if (c.lb_flags[i] || o.flags) {
      *w++ = 0;
      continue;
    }
    const uint64_t a = c.lb_tok[i], n = c.lb_tok[i + 1] - a, N = o.nodes;
    *w++ = 1;
    memcpy(w, &n, 8);
    w += 8;
    memcpy(w, v[0].h + a * tb, n * tb);
    w += n * tb;
    if (ib == 8) {
      memcpy(w, v[1].h + a * 8, n * 8);
      w += n * 8;
    } else {
      for (uint64_t k = 0; k < n; ++k, w += ib) memcpy(w, v[1].h + (a + k) * 8, ib);
    }
    memcpy(w, v[2].h + a * lb, n * lb);
    w += n * lb;
    memcpy(w, &N, 8);
    w += 8;
    memcpy(w, v[3].h + o.node_off * pb, N * pb);
    w += N * pb;
    if (ib == 8) {
      memcpy(w, v[4].h + o.node_off * 8, N * 8);
      w += N * 8;
      memcpy(w, v[5].h + o.term_off * 8, n * 8);
      w += n * 8;
    } else {
      for (uint64_t k = 0; k < N; ++k, w += ib) memcpy(w, v[4].h + (o.node_off + k) * 8, ib);
      for (uint64_t k = 0; k < n; ++k, w += ib) memcpy(w, v[5].h + (o.term_off + k) * 8, ib);
    }
  }
  return 0;
}

namespace {

// Records appended to one malloc'd buffer.
struct Records {
  uint8_t* buf = nullptr;
  size_t len = 0, cap = 0;
  int append(const uint8_t* p, size_t n) {
    if (len + n > cap) {
      size_t nc = cap ? cap : 1 << 16;
      while (nc < len + n) nc *= 2;
      uint8_t* q = static_cast<uint8_t*>(realloc(buf, nc));
      if (!q) return fail(ALPACC_ERR_NOMEM, "host allocation failed");
      buf = q;
      cap = nc;
    }
    memcpy(buf + len, p, n);
    len += n;
    return 0;
  }
};

// Inputs [a, b) of a batch, answered in launches that each stay within one
// launch's limits: at most BATCH_LAUNCH_BYTES of input (the lexer's 31-bit
// positions), and in parser mode at most parse_max_positions framed
// positions.  In combined mode the token count is known only after the
// lexer, so a launch the parser refuses for its size is split in halves.
// An input beyond BATCH_LAUNCH_BYTES alone is answered by run_case: past the
// lexer's launch limit through the stream path (in combined mode followed
// by one parser launch over all its tokens); in parser mode by one parser
// launch, refused beyond that launch's positions.
constexpr uint64_t BATCH_LAUNCH_BYTES = 1ull << 30;

int records_in_launches(Context& c, int mode, const uint8_t* packed, const uint64_t* off, uint64_t a, uint64_t b,
                        Records& acc) {
  const uint64_t max_pos = parse_max_positions(c.g);
  uint64_t i = a;
  while (i < b) {
    if (off[i + 1] - off[i] > BATCH_LAUNCH_BYTES) {  // one input of its own
      std::vector<uint8_t> rec;
      int r;
      switch (mode) {
        case ALPACC_MODE_LEXER: r = lexer_record(c, off[i + 1] - off[i], packed + off[i], rec); break;
        case ALPACC_MODE_PARSER: r = parser_record(c, off[i + 1] - off[i], packed + off[i], rec); break;
        default: r = both_record(c, off[i + 1] - off[i], packed + off[i], rec); break;
      }
      if (r < 0) return r;
      if ((r = acc.append(rec.data(), rec.size()))) return r;
      ++i;
      continue;
    }
    uint64_t j = i + 1;
    auto fits = [&](uint64_t e) {
      if (off[e] - off[i] > BATCH_LAUNCH_BYTES) return false;
      return mode != ALPACC_MODE_PARSER || (off[e] - off[i]) / c.g->terminal_bytes_both + 2 * (e - i) <= max_pos;
    };
    while (j < b && fits(j + 1)) ++j;
    uint8_t* out = nullptr;
    size_t len = 0;
    int r = batch_records(c, mode, packed, off + i, j - i, &out, &len, nullptr);
    if (r == ALPACC_ERR_LIMIT && j - i > 1) {  // too many tokens for one parser launch
      free(out);
      const uint64_t mid = i + (j - i) / 2;
      if ((r = records_in_launches(c, mode, packed, off, i, mid, acc)) < 0) return r;
      if ((r = records_in_launches(c, mode, packed, off, mid, j, acc)) < 0) return r;
      i = j;
      continue;
    }
    if (r < 0) {
      free(out);
      return r;
    }
    if (!acc.buf) {  // the first launch's records become the buffer
      acc.buf = out;
      acc.len = acc.cap = len;
    } else {
      r = acc.append(out, len);
      free(out);
      if (r) return r;
    }
    i = j;
  }
  return 0;
}

}  // namespace

}  // namespace alpacc

using namespace alpacc;

extern "C" {

int alpacc_hip_run_batch(alpacc_hip_ctx* h, int mode, const uint8_t* packed, const uint64_t* offsets, uint64_t count,
                         uint8_t** records, uint64_t* records_len) {
  if (!h || !offsets || !records || !records_len || (!packed && count && offsets[count] > offsets[0]))
    return fail(ALPACC_ERR_ARG, "null argument");
  if (mode != ALPACC_MODE_LEXER && mode != ALPACC_MODE_PARSER && mode != ALPACC_MODE_BOTH)
    return fail(ALPACC_ERR_ARG, "bad mode %d", mode);
  *records = nullptr;
  *records_len = 0;
  if (count == 0) return 0;
  for (uint64_t i = 0; i < count; ++i)
    if (offsets[i + 1] < offsets[i]) return fail(ALPACC_ERR_ARG, "offsets decrease at input %llu", (unsigned long long)i);
  Context& c = h->c;
  HIP_TRY(hipSetDevice(c.device));
  Records acc;
  int r = records_in_launches(c, mode, packed, offsets, 0, count, acc);
  if (r < 0) {
    free(acc.buf);
    return r;
  }
  *records = acc.buf;
  *records_len = acc.len;
  return 0;
}

int alpacc_hip_set_graphs(alpacc_hip_ctx* h, int on) {
  if (!h) return fail(ALPACC_ERR_ARG, "null argument");
  Context& c = h->c;
  c.use_graphs = on != 0;
  if (!c.use_graphs && c.graph) {
    HIP_TRY(hipSetDevice(c.device));
    (void)hipGraphExecDestroy(c.graph);
    c.graph = nullptr;
  }
  return 0;
}

int alpacc_hip_set_bracket_width(alpacc_hip_ctx* h, int bytes) {
  if (!h) return fail(ALPACC_ERR_ARG, "null argument");
  if (bytes < 0 || bytes > 2) return fail(ALPACC_ERR_ARG, "bracket width %d is not 0, 1 or 2", bytes);
  Context& c = h->c;
  if (bytes == 1 && bracket_symbol_max(c.g) >= 256)
    return fail(ALPACC_ERR_ARG, "bracket symbols of %s reach 256: one-byte stacks cannot hold them", c.g->name);
  HIP_TRY(hipSetDevice(c.device));
  c.resolve_bw = bytes;
  if (c.graph) {  // captured with the old kernels
    (void)hipGraphExecDestroy(c.graph);
    c.graph = nullptr;
  }
  return 0;
}

int alpacc_hip_graph_stats(const alpacc_hip_ctx* h, uint64_t* captures, uint64_t* replays) {
  if (!h || !captures || !replays) return fail(ALPACC_ERR_ARG, "null argument");
  *captures = h->c.graph_captures;
  *replays = h->c.graph_replays;
  return 0;
}

int alpacc_hip_lex_parse_batch_device(alpacc_hip_ctx* h, const uint8_t* d_packed, const uint64_t* offsets,
                                      uint64_t count, alpacc_hip_batch* out, void* stream) {
  if (!h || !offsets || !out || count == 0) return fail(ALPACC_ERR_ARG, "null argument or empty batch");
  Context& c = h->c;
  HIP_TRY(hipSetDevice(c.device));
  memset(out, 0, sizeof *out);
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c.stream;
  int r = lex_parse_batch(c, d_packed, offsets, count, st);
  if (r < 0) return r;
  c.bh_valid.resize(count);
  c.bh_node_off.resize(count);
  c.bh_nodes.resize(count);
  c.bh_tp_off.resize(count);
  for (uint64_t i = 0; i < count; ++i) {
    const ParseResult& o = c.pb_res[i];
    const bool ok = !c.lb_flags[i] && !o.flags;
    c.bh_valid[i] = ok;
    c.bh_node_off[i] = ok ? o.node_off : 0;
    c.bh_nodes[i] = ok ? o.nodes : 0;
    c.bh_tp_off[i] = ok ? o.term_off : 0;
  }
  out->count = count;
  out->valid = c.bh_valid.data();
  out->token_offsets = c.lb_tok;
  out->node_offsets = c.bh_node_off.data();
  out->node_counts = c.bh_nodes.data();
  out->token_parent_offsets = c.bh_tp_off.data();
  out->tokens.count = c.lb_tok[count];
  out->tokens.terminals = c.lx_term.ptr;
  out->tokens.starts = c.lx_start.as<int64_t>();
  out->tokens.lengths = c.lx_len.ptr;
  out->productions = c.tr_prod.ptr;
  out->parents = c.tr_parent.as<int64_t>();
  out->token_parents = c.tr_tokpar.as<int64_t>();
  return 0;
}

}  // extern "C"
