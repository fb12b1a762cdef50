// Parser and combined entry points of the C ABI (parse_test, the combined
// lex + parse + tree run, their device forms and wire records).
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
int lexemes_to_host(Context& c, int mode, uint64_t count, hipStream_t st, alpacc_hip_lexemes* out);
void lex_stream_begin(Context& c, int mode);
int lex_stream_push(Context& c, const uint8_t* d_chunk, uint64_t n, int is_last_chunk, alpacc_hip_lexemes* d_out,
                    hipStream_t st);

static int lex_parse_device(Context& c, const uint8_t* d_in, uint64_t n, hipStream_t st, alpacc_hip_tree* out) {
  uint64_t count = 0;
  int r = lex_on_device(c, ALPACC_MODE_BOTH, d_in, n, st, &count);
  if (r != ALPACC_VALID) return r;
  uint64_t np = 0;
  r = parse_on_device(c, c.lx_term.as<uint8_t>(), count, true, st, &np);
  if (r != ALPACC_VALID) return r;
  out->tokens.count = count;
  out->tokens.terminals = c.lx_term.ptr;
  out->tokens.starts = c.lx_start.as<int64_t>();
  out->tokens.lengths = c.lx_len.ptr;
  out->num_nodes = np - count;
  out->productions = c.tr_prod.ptr;
  out->parents = c.tr_parent.as<int64_t>();
  out->token_parents = c.tr_tokpar.as<int64_t>();
  return ALPACC_VALID;
}

static void* host_copy(const void* d, size_t bytes, hipStream_t st, int* err) {
  void* h = malloc(bytes ? bytes : 1);
  if (!h) {
    *err = fail(ALPACC_ERR_NOMEM, "host allocation failed");
    return nullptr;
  }
  if (bytes) {
    hipError_t e = hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) {
      free(h);
      *err = fail(ALPACC_ERR_HIP, "hipMemcpyAsync: %s", hipGetErrorString(e));
      return nullptr;
    }
  }
  return h;
}

// Parser-only record: u8 1, u64 count, production_t ids (C/Generator.hs:64-84).
int parser_record(Context& c, uint64_t n, const uint8_t* payload, std::vector<uint8_t>& rec) {
  hipStream_t st = c.stream;
  const uint8_t* d_tok = nullptr;
  int r;
  if ((r = stage_input(c, payload, n, st, &d_tok))) return r;
  uint64_t np = 0;
  r = parse_on_device(c, d_tok, n, false, st, &np);
  if (r < 0) return r;
  if (r == ALPACC_REJECT) {
    rec.assign(1, 0);
    return r;
  }
  const uint64_t pb = (uint64_t)c.g->production_bytes;
  rec.resize(9 + np * pb);
  rec[0] = 1;
  memcpy(&rec[1], &np, 8);
  if (np) HIP_TRY(hipMemcpyAsync(rec.data() + 9, c.ps_d.ptr, np * pb, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return ALPACC_VALID;
}

// Combined record of an input beyond one lexer launch: its bytes are lexed
// in chunks of LEX_STREAM_CHUNK through the stream path (combined
// numbering; the tokens gathered on the host as they close), then its
// terminals are parsed in one launch.  An input whose tokens pass one
// parser launch's positions is refused (a size error).
static int both_record_stream(Context& c, uint64_t n, const uint8_t* payload, std::vector<uint8_t>& rec) {
  hipStream_t st = c.stream;
  const GrammarDesc* g = c.g;
  const int tb = g->terminal_bytes_both, ib = g->index_bytes, lb = g->length_bytes, pb = g->production_bytes;
  std::vector<uint8_t> ht, hl;
  std::vector<int64_t> hs;
  uint64_t T = 0;
  int r;
  lex_stream_begin(c, ALPACC_MODE_BOTH);
  for (uint64_t at = 0; at < n;) {
    const uint64_t m = std::min(n - at, LEX_STREAM_CHUNK);
    const bool last = at + m == n;
    const uint8_t* d_in = nullptr;
    alpacc_hip_lexemes d{};
    if ((r = stage_input(c, payload + at, m, st, &d_in)) || (r = lex_stream_push(c, d_in, m, last ? 1 : 0, &d, st)) < 0) {
      c.lstream.on = false;
      return r;
    }
    if (r == ALPACC_REJECT) {
      c.lstream.on = false;
      rec.assign(1, 0);
      return ALPACC_REJECT;
    }
    if (d.count) {  // the chunk's tokens, starts counted from the input's first byte
      ht.resize((T + d.count) * tb);
      hs.resize(T + d.count);
      hl.resize((T + d.count) * lb);
      HIP_TRY(hipMemcpyAsync(ht.data() + T * tb, d.terminals, d.count * tb, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipMemcpyAsync(hs.data() + T, d.starts, d.count * 8, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipMemcpyAsync(hl.data() + T * lb, d.lengths, d.count * lb, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      T += d.count;
    }
    at += m;
  }
  if (T + 2 > parse_max_positions(g))
    return fail(ALPACC_ERR_LIMIT, "%llu tokens exceed one parser launch (%llu framed positions)", (unsigned long long)T,
                (unsigned long long)parse_max_positions(g));
  const uint8_t* d_tok = nullptr;
  if ((r = stage_input(c, ht.data(), T * tb, st, &d_tok))) return r;
  uint64_t np = 0;
  if ((r = parse_on_device(c, d_tok, T, true, st, &np)) < 0) return r;
  if (r == ALPACC_REJECT) {
    rec.assign(1, 0);
    return r;
  }
  const uint64_t N = np - T;
  rec.resize(1 + 8 + T * (tb + ib + lb) + 8 + N * (pb + ib) + T * ib);
  uint8_t* o = rec.data();
  *o++ = 1;
  memcpy(o, &T, 8);
  o += 8;
  // indices are i64 here and on the device; a 32-bit index_t keeps the low
  // half of each (little-endian)
  auto put_index = [&](const int64_t* v, uint64_t count) {
    if (ib == 8) {
      memcpy(o, v, count * 8);
    } else {
      for (uint64_t i = 0; i < count; ++i) memcpy(o + i * ib, &v[i], ib);
    }
    o += count * ib;
  };
  memcpy(o, ht.data(), T * tb);
  o += T * tb;
  put_index(hs.data(), T);
  memcpy(o, hl.data(), T * lb);
  o += T * lb;
  memcpy(o, &N, 8);
  o += 8;
  if (N) HIP_TRY(hipMemcpyAsync(o, c.tr_prod.ptr, N * pb, hipMemcpyDeviceToHost, st));
  o += N * pb;
  std::vector<int64_t> par(std::max(N, T));
  if (N) HIP_TRY(hipMemcpyAsync(par.data(), c.tr_parent.ptr, N * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  put_index(par.data(), N);
  if (T) HIP_TRY(hipMemcpyAsync(par.data(), c.tr_tokpar.ptr, T * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  put_index(par.data(), T);
  return ALPACC_VALID;
}

// Combined record (C/Generator.hs:114-152): SoA tokens, then the compact tree.
int both_record(Context& c, uint64_t n, const uint8_t* payload, std::vector<uint8_t>& rec) {
  if (n > LEX_MAX_LAUNCH_BYTES) return both_record_stream(c, n, payload, rec);
  hipStream_t st = c.stream;
  const uint8_t* d_in = nullptr;
  int r;
  if ((r = stage_input(c, payload, n, st, &d_in))) return r;
  alpacc_hip_tree t{};
  r = lex_parse_device(c, d_in, n, st, &t);
  if (r < 0) return r;
  if (r == ALPACC_REJECT) {
    rec.assign(1, 0);
    return r;
  }
  const GrammarDesc* g = c.g;
  const uint64_t T = t.tokens.count, N = t.num_nodes;
  const int tb = g->terminal_bytes_both, ib = g->index_bytes, lb = g->length_bytes, pb = g->production_bytes;
  rec.resize(1 + 8 + T * (tb + ib + lb) + 8 + N * (pb + ib) + T * ib);
  uint8_t* o = rec.data();
  *o++ = 1;
  memcpy(o, &T, 8);
  o += 8;
  auto pull = [&](const void* d, size_t bytes) -> int {
    if (bytes) HIP_TRY(hipMemcpyAsync(o, d, bytes, hipMemcpyDeviceToHost, st));
    o += bytes;
    return 0;
  };
  // i64 on the device; a 32-bit index_t keeps the low half of each (values
  // stay below 2^31: one launch takes fewer than 2^31 bytes)
  std::vector<int64_t> wide;
  struct Narrow {
    uint8_t* at;
    size_t from, count;
  };
  std::vector<Narrow> narrow;
  auto pull_index = [&](const int64_t* d, uint64_t count) -> int {
    if (ib == 8) return pull(d, count * 8);
    const size_t from = wide.size();
    wide.resize(from + count);
    if (count) HIP_TRY(hipMemcpyAsync(wide.data() + from, d, count * 8, hipMemcpyDeviceToHost, st));
    narrow.push_back({o, from, count});
    o += count * ib;
    return 0;
  };
  wide.reserve(ib == 8 ? 0 : (size_t)(2 * T + N));
  if ((r = pull(t.tokens.terminals, T * tb)) || (r = pull_index(t.tokens.starts, T)) ||
      (r = pull(t.tokens.lengths, T * lb)))
    return r;
  memcpy(o, &N, 8);
  o += 8;
  if ((r = pull(t.productions, N * pb)) || (r = pull_index(t.parents, N)) || (r = pull_index(t.token_parents, T)))
    return r;
  HIP_TRY(hipStreamSynchronize(st));
  for (const Narrow& x : narrow)
    for (size_t k = 0; k < x.count; ++k) {
      const int32_t v = (int32_t)wide[x.from + k];
      memcpy(x.at + k * ib, &v, ib);
    }
  return ALPACC_VALID;
}

}  // namespace alpacc

using namespace alpacc;

extern "C" {

int alpacc_hip_parse_device(alpacc_hip_ctx* h, const void* d_tokens, uint64_t n, void** d_prods, uint64_t* num_prods,
                            void* stream) {
  if (!h || !d_prods || !num_prods || (n && !d_tokens)) return fail(ALPACC_ERR_ARG, "null argument");
  Context& c = h->c;
  HIP_TRY(hipSetDevice(c.device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c.stream;
  int r = parse_on_device(c, static_cast<const uint8_t*>(d_tokens), n, false, st, num_prods);
  *d_prods = r == ALPACC_VALID ? c.ps_d.ptr : nullptr;
  return r;
}

int alpacc_hip_parse_test(alpacc_hip_ctx* h, const void* tokens, uint64_t n, void** prods, uint64_t* num_prods) {
  if (!h || !prods || !num_prods || (n && !tokens)) return fail(ALPACC_ERR_ARG, "null argument");
  Context& c = h->c;
  HIP_TRY(hipSetDevice(c.device));
  *prods = nullptr;
  *num_prods = 0;
  const uint8_t* d_tok = nullptr;
  int r;
  if ((r = stage_input(c, tokens, n * c.g->terminal_bytes_both, c.stream, &d_tok))) return r;
  uint64_t np = 0;
  r = parse_on_device(c, d_tok, n, false, c.stream, &np);
  if (r != ALPACC_VALID) return r;
  int err = 0;
  void* hp = host_copy(c.ps_d.ptr, np * c.g->production_bytes, c.stream, &err);
  if (!hp) return err;
  HIP_TRY(hipStreamSynchronize(c.stream));
  *prods = hp;
  *num_prods = np;
  return ALPACC_VALID;
}

int alpacc_hip_lex_parse_device(alpacc_hip_ctx* h, const uint8_t* d_str, uint64_t n, alpacc_hip_tree* d_out,
                                void* stream) {
  if (!h || !d_out || (n && !d_str)) return fail(ALPACC_ERR_ARG, "null argument");
  Context& c = h->c;
  HIP_TRY(hipSetDevice(c.device));
  memset(d_out, 0, sizeof *d_out);
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c.stream;
  return lex_parse_device(c, d_str, n, st, d_out);
}

int alpacc_hip_lex_parse(alpacc_hip_ctx* h, const uint8_t* str, uint64_t n, alpacc_hip_tree* out) {
  if (!h || !out || (n && !str)) return fail(ALPACC_ERR_ARG, "null argument");
  Context& c = h->c;
  HIP_TRY(hipSetDevice(c.device));
  memset(out, 0, sizeof *out);
  const uint8_t* d_in = nullptr;
  int r;
  if ((r = stage_input(c, str, n, c.stream, &d_in))) return r;
  alpacc_hip_tree t{};
  r = lex_parse_device(c, d_in, n, c.stream, &t);
  if (r != ALPACC_VALID) return r;
  const GrammarDesc* g = c.g;
  const uint64_t T = t.tokens.count, N = t.num_nodes;
  int err = 0;
  out->tokens.count = T;
  out->num_nodes = N;
  out->tokens.terminals = host_copy(t.tokens.terminals, T * g->terminal_bytes_both, c.stream, &err);
  out->tokens.starts = static_cast<int64_t*>(host_copy(t.tokens.starts, T * 8, c.stream, &err));
  out->tokens.lengths = host_copy(t.tokens.lengths, T * g->length_bytes, c.stream, &err);
  out->productions = host_copy(t.productions, N * g->production_bytes, c.stream, &err);
  out->parents = static_cast<int64_t*>(host_copy(t.parents, N * 8, c.stream, &err));
  out->token_parents = static_cast<int64_t*>(host_copy(t.token_parents, T * 8, c.stream, &err));
  HIP_TRY(hipStreamSynchronize(c.stream));
  if (err) return err;
  return ALPACC_VALID;
}

}  // extern "C"

// ---- parser tile shape and its autotuner --------------------------------

extern "C" {

int alpacc_hip_set_parse_shape(alpacc_hip_ctx* h, const alpacc_hip_parse_shape* shape) {
  if (!h || !shape) return fail(ALPACC_ERR_ARG, "null argument");
  if (!parse_shape_ok(shape->positions_per_tile, shape->positions_per_lane))
    return fail(ALPACC_ERR_ARG, "parser tile shape %d positions per tile, %d per lane is not one of the kernels'",
                shape->positions_per_tile, shape->positions_per_lane);
  Context& c = h->c;
  HIP_TRY(hipSetDevice(c.device));
  c.parse_pt = shape->positions_per_tile;
  c.parse_apl = shape->positions_per_lane;
  c.cap_stage_n = c.cap_stage_t = 0;  // staging learned for other tiles
  c.cap_tokens = 0;
  if (c.graph) {  // captured with the old kernels
    (void)hipGraphExecDestroy(c.graph);
    c.graph = nullptr;
  }
  return 0;
}

int alpacc_hip_get_parse_shape(const alpacc_hip_ctx* h, alpacc_hip_parse_shape* shape) {
  if (!h || !shape) return fail(ALPACC_ERR_ARG, "null argument");
  shape->positions_per_tile = h->c.parse_pt;
  shape->positions_per_lane = h->c.parse_apl;
  shape->ms = 0;
  return 0;
}

int alpacc_hip_autotune_parser(alpacc_hip_ctx* h, const uint8_t* d_sample, uint64_t n, int reps,
                               alpacc_hip_parse_shape* best) {
  if (!h || !best || !d_sample || n == 0 || reps <= 0) return fail(ALPACC_ERR_ARG, "null argument or empty sample");
  Context& c = h->c;
  HIP_TRY(hipSetDevice(c.device));
  if (!c.g->has_parser) return fail(ALPACC_ERR_ARG, "grammar '%s' has no production rules", c.g->name);
  for (hipEvent_t& ev : c.ev)
    if (!ev) HIP_TRY(hipEventCreate(&ev));
  const bool timing = c.timing, ktiming = c.timer.on;
  c.timing = true;  // the parser's stage events time each run
  c.timer.on = false;
  static const int shapes[][2] = {{1024, 4}, {512, 4}, {2048, 4}, {1024, 2}, {1024, 8}};
  alpacc_hip_parse_shape top{c.parse_pt, c.parse_apl, 0};
  double top_ms = -1;
  int r = 0;
  for (const auto& sh : shapes) {
    const alpacc_hip_parse_shape want{sh[0], sh[1], 0};
    if ((r = alpacc_hip_set_parse_shape(h, &want)) < 0) break;
    alpacc_hip_tree t{};
    if ((r = lex_parse_device(c, d_sample, n, c.stream, &t)) < 0) break;  // warm
    c.parse_ms = 0;
    c.parse_launches = 0;
    for (int i = 0; i < reps && r >= 0; ++i) r = lex_parse_device(c, d_sample, n, c.stream, &t);
    if (r < 0) break;
    if (!c.parse_launches) {
      r = fail(ALPACC_ERR_ARG, "the tuning sample is rejected by the lexer: the parser never ran");
      break;
    }
    const double ms = c.parse_ms / (double)c.parse_launches;
    if (top_ms < 0 || ms < top_ms) {
      top_ms = ms;
      top = want;
    }
  }
  c.timing = timing;
  c.timer.on = ktiming;
  c.parse_ms = c.lex_ms = 0;
  c.parse_launches = c.lex_launches = 0;
  if (r < 0) return r;
  top.ms = top_ms;
  if ((r = alpacc_hip_set_parse_shape(h, &top)) < 0) return r;
  *best = top;
  return 0;
}

}  // extern "C"
