// Host runtime of the HIP backend: grammar contexts with device-resident
// tables, grow-only device workspaces, and the C ABI of include/alpacc_hip.h.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/alpacc_hip.h"
#include "grammar.hpp"
#include "lexer.hpp"
#include "parser.hpp"
#include "runtime.hpp"

namespace alpacc {

static thread_local std::string g_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return fail(ALPACC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

static void free_retired(std::vector<void*>& v, hipError_t (*f)(void*)) {
  for (void* p : v) (void)f(p);
  v.clear();
}

int DevBuf::reserve(size_t want) {
  if (want <= cap) return 0;
  const size_t old = cap;
  if (ptr) retired.push_back(ptr);
  ptr = nullptr;
  cap = 0;
  // at least double: a workspace that keeps growing (a server's batches)
  // is reallocated a logarithmic number of times
  size_t sz = want < 2 * old ? 2 * old : want;
  sz = sz < 256 ? 256 : sz;
  sz = (sz + 255) & ~size_t(255);
  hipError_t e = hipMalloc(&ptr, sz);
  if (e != hipSuccess && !retired.empty()) {  // room from the retired blocks (this waits for the device)
    (void)hipGetLastError();
    free_retired(retired, hipFree);
    e = hipMalloc(&ptr, sz);
  }
  if (e != hipSuccess && sz > want) {  // no room to double: the exact size
    (void)hipGetLastError();
    sz = (want + 255) & ~size_t(255);
    e = hipMalloc(&ptr, sz);
  }
  if (e != hipSuccess) {
    ptr = nullptr;
    return fail(ALPACC_ERR_NOMEM, "hipMalloc(%zu): %s", sz, hipGetErrorString(e));
  }
  cap = sz;
  ++gen;
  return 0;
}

void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  free_retired(retired, hipFree);
  ptr = nullptr;
  cap = 0;
}

int PinnedBuf::reserve(size_t want) {
  if (want <= cap) return 0;
  const size_t old = cap;
  if (ptr) retired.push_back(ptr);
  ptr = nullptr;
  cap = 0;
  size_t sz = want < 2 * old ? 2 * old : want;  // at least double, as DevBuf
  sz = sz < 4096 ? 4096 : sz;
  sz = (sz + 4095) & ~size_t(4095);
  hipError_t e = hipHostMalloc(&ptr, sz, hipHostMallocDefault);
  if (e != hipSuccess && !retired.empty()) {
    (void)hipGetLastError();
    free_retired(retired, hipHostFree);
    e = hipHostMalloc(&ptr, sz, hipHostMallocDefault);
  }
  if (e != hipSuccess && sz > want) {
    (void)hipGetLastError();
    sz = (want + 4095) & ~size_t(4095);
    e = hipHostMalloc(&ptr, sz, hipHostMallocDefault);
  }
  if (e != hipSuccess) {
    ptr = nullptr;
    return fail(ALPACC_ERR_NOMEM, "hipHostMalloc(%zu): %s", sz, hipGetErrorString(e));
  }
  cap = sz;
  ++gen;
  return 0;
}

void PinnedBuf::release() {
  if (ptr) (void)hipHostFree(ptr);
  free_retired(retired, hipHostFree);
  ptr = nullptr;
  cap = 0;
}

template <typename T>
static int upload(T** dst, const T* src, size_t count) {
  size_t bytes = (count ? count : 1) * sizeof(T);
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(dst), bytes));
  HIP_TRY(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// Zero-fill of lookback state by a kernel (no memset nodes, so the launch
// sequence stays safe to capture in a graph).
__global__ void fill_zero_u32(uint32_t* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0;
}

int zero_words(void* p, size_t bytes, hipStream_t st) {
  size_t n = (bytes + 3) / 4;
  int blocks = (int)((n + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(fill_zero_u32, dim3(blocks), dim3(256), 0, st, static_cast<uint32_t*>(p), n);
  HIP_TRY(hipGetLastError());
  return 0;
}

// Several ranges in one launch: workgroup row y clears range y.
struct ZeroRanges {
  uint32_t* p[ZERO_RANGES_MAX];
  size_t n[ZERO_RANGES_MAX];
};
__global__ void fill_zero_ranges(ZeroRanges z) {
  uint32_t* p = z.p[blockIdx.y];
  const size_t n = z.n[blockIdx.y];
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

int zero_ranges(std::initializer_list<std::pair<void*, size_t>> ranges, hipStream_t st) {
  if (ranges.size() < 1 || ranges.size() > ZERO_RANGES_MAX) return fail(ALPACC_ERR_ARG, "zero_ranges of %zu ranges", ranges.size());
  ZeroRanges z{};
  size_t most = 0;
  int k = 0;
  for (const auto& r : ranges) {
    z.p[k] = static_cast<uint32_t*>(r.first);
    z.n[k] = (r.second + 3) / 4;
    most = z.n[k] > most ? z.n[k] : most;
    ++k;
  }
  int blocks = (int)((most + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(fill_zero_ranges, dim3(blocks, k), dim3(256), 0, st, z);
  HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace alpacc

using namespace alpacc;

static int check_mode(int mode) {
  if (mode < ALPACC_MODE_LEXER || mode > ALPACC_MODE_BOTH) return fail(ALPACC_ERR_ARG, "bad mode %d", mode);
  return 0;
}

extern "C" {

const char* alpacc_hip_last_error(void) { return g_error.c_str(); }

int alpacc_hip_open(const char* grammar, int device, alpacc_hip_ctx** out) {
  if (!grammar || !out) return fail(ALPACC_ERR_ARG, "null argument");
  *out = nullptr;
  const GrammarDesc* g = find_grammar(grammar);
  if (!g) return fail(ALPACC_ERR_ARG, "unknown grammar '%s'", grammar);
  HIP_TRY(hipSetDevice(device));
  auto* h = new alpacc_hip_ctx();
  Context& c = h->c;
  c.g = g;
  c.device = device;
  int r = 0;
  hipError_t e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete h;
    return fail(ALPACC_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  if ((r = upload(&c.d_class, g->byte_class, 256)) || (r = upload(&c.d_step, g->step, (size_t)g->nendo * g->nclass)) ||
      (r = upload(&c.d_compose, g->compose, (size_t)g->nendo * g->nendo)) ||
      (r = upload(&c.d_endo_init, g->endo_init, (size_t)g->nendo)) ||
      (r = upload(&c.d_delta_lex, g->delta_lex, (size_t)g->nstate * 256)) ||
      (r = upload(&c.d_delta_both, g->delta_both, (size_t)g->nstate * 256)) ||
      (r = upload(&c.d_token_lex, g->token_lex, (size_t)g->nstate)) ||
      (r = upload(&c.d_token_both, g->token_both, (size_t)g->nstate)) ||
      (r = upload(&c.d_pair_rec, g->pair_rec, (size_t)(g->npairs + 1) * 64)) ||
      (g->production_bytes == 2 && (r = upload(&c.d_prod16, g->pair_prod16, (size_t)(g->npairs + 1) * 8))) ||
      (r = c.small.reserve(4096))) {
    alpacc_hip_close(h);
    return r;
  }
  {  // window entries: pair id and the first 6 bytes of its record
    std::vector<uint64_t> wr((size_t)g->nwindow);
    for (size_t i = 0; i < wr.size(); ++i) {
      const uint32_t id = g->window[i];
      uint64_t v = id;
      if (id != 0xFFFFu)
        for (int b = 0; b < 6; ++b) v |= (uint64_t)g->pair_rec[(size_t)id * 64 + b] << (16 + 8 * b);
      wr[i] = v;
    }
    if ((r = upload(&c.d_wrec, wr.data(), wr.size())) ||
        (g->window_hash && (r = upload(&c.d_wkeys, g->window_keys, (size_t)g->nwindow)))) {
      alpacc_hip_close(h);
      return r;
    }
  }
  e = hipHostMalloc(reinterpret_cast<void**>(&c.host_small), 4096, hipHostMallocDefault);
  if (e != hipSuccess) {
    c.host_small = nullptr;
    alpacc_hip_close(h);
    return fail(ALPACC_ERR_NOMEM, "hipHostMalloc: %s", hipGetErrorString(e));
  }
  *out = h;
  return 0;
}

void alpacc_hip_close(alpacc_hip_ctx* h) {
  if (!h) return;
  Context& c = h->c;
  (void)hipSetDevice(c.device);
  if (c.graph) (void)hipGraphExecDestroy(c.graph);
  void* ptrs[] = {c.d_class, c.d_step, c.d_compose, c.d_endo_init, c.d_delta_lex, c.d_delta_both, c.d_token_lex, c.d_token_both, c.d_wrec, c.d_wkeys,
                  c.d_pair_rec, c.d_prod16};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (DevBuf* b : c.all_buffers()) b->release();
  if (c.host_small) (void)hipHostFree(c.host_small);
  for (hipEvent_t ev : c.ev)
    if (ev) (void)hipEventDestroy(ev);
  c.timer.destroy();
  c.h_meta.release();
  c.h_pmeta.release();
  c.h_pres.release();
  c.h_out.release();
  if (c.stream) (void)hipStreamDestroy(c.stream);
  delete h;
}

static int layout_text(const GrammarDesc* g, int mode, char* buf, size_t cap) {
  if (mode != ALPACC_MODE_LEXER && !g->has_parser)
    return fail(ALPACC_ERR_ARG, "grammar '%s' has no production rules: lexer mode only", g->name);
  std::string s;
  char line[64];
  int tb = mode == ALPACC_MODE_LEXER ? g->terminal_bytes_lex : g->terminal_bytes_both;
  snprintf(line, sizeof line, "terminal_t=%d\n", tb);
  s += line;
  if (mode != ALPACC_MODE_LEXER) {
    snprintf(line, sizeof line, "production_t=%d\n", g->production_bytes);
    s += line;
  }
  snprintf(line, sizeof line, "index_t=%d\n", g->index_bytes);
  s += line;
  if (mode != ALPACC_MODE_PARSER) {
    snprintf(line, sizeof line, "length_t=%d\n", g->length_bytes);
    s += line;
  }
  snprintf(line, sizeof line, "workgroup_size=%d\nitems_per_thread=%d\n", LEX_WG, LEX_IPT);
  s += line;
  if (s.size() + 1 > cap) return fail(ALPACC_ERR_ARG, "layout buffer too small");
  memcpy(buf, s.c_str(), s.size() + 1);
  return (int)s.size();
}

int alpacc_hip_grammar_layout(const char* grammar, int mode, char* buf, size_t cap) {
  if (!grammar || !buf) return fail(ALPACC_ERR_ARG, "null argument");
  if (int r = check_mode(mode)) return r;
  const GrammarDesc* g = find_grammar(grammar);
  if (!g) return fail(ALPACC_ERR_ARG, "unknown grammar '%s'", grammar);
  return layout_text(g, mode, buf, cap);
}

int alpacc_hip_layout(const alpacc_hip_ctx* h, int mode, char* buf, size_t cap) {
  if (!h || !buf) return fail(ALPACC_ERR_ARG, "null argument");
  if (int r = check_mode(mode)) return r;
  return layout_text(h->c.g, mode, buf, cap);  // the context's widths
}

// The descriptor of a width variant (see alpacc_hip_open_variant).
static int variant_desc(const GrammarDesc* g, int index32, int length_bits, GrammarDesc* v) {
  if (length_bits != 0 && length_bits != 8 && length_bits != 16 && length_bits != 32 && length_bits != 64)
    return fail(ALPACC_ERR_ARG, "length %d is not 8, 16, 32 or 64", length_bits);
  *v = *g;
  v->index_bytes = index32 ? 4 : 8;
  // length_t: the requested width, else the grammar's `length`, else the
  // unsigned counterpart of index_t
  const int bits = length_bits ? length_bits : v->length_param;
  v->length_bytes = bits ? bits / 8 : v->index_bytes;
  return 0;
}

int alpacc_hip_variant_layout(const char* grammar, int mode, int index32, int length_bits, char* buf, size_t cap) {
  if (!grammar || !buf) return fail(ALPACC_ERR_ARG, "null argument");
  if (int r = check_mode(mode)) return r;
  const GrammarDesc* g = find_grammar(grammar);
  if (!g) return fail(ALPACC_ERR_ARG, "unknown grammar '%s'", grammar);
  GrammarDesc v;
  if (int r = variant_desc(g, index32, length_bits, &v)) return r;
  return layout_text(&v, mode, buf, cap);
}

int alpacc_hip_open_variant(const char* grammar, int device, int index32, int length_bits, alpacc_hip_ctx** out) {
  if (!out || !grammar) return fail(ALPACC_ERR_ARG, "null argument");
  *out = nullptr;
  const GrammarDesc* g = find_grammar(grammar);
  if (!g) return fail(ALPACC_ERR_ARG, "unknown grammar '%s'", grammar);
  GrammarDesc v;
  if (int r = variant_desc(g, index32, length_bits, &v)) return r;
  alpacc_hip_ctx* h = nullptr;
  int r = alpacc_hip_open(grammar, device, &h);
  if (r < 0) return r;
  h->c.gv = v;
  h->c.g = &h->c.gv;
  *out = h;
  return 0;
}

int alpacc_hip_grammar_limits(const char* grammar, uint64_t* launch_bytes, uint64_t* parser_positions) {
  const GrammarDesc* g = grammar ? find_grammar(grammar) : nullptr;
  if (!g || !launch_bytes || !parser_positions) return fail(ALPACC_ERR_ARG, "unknown grammar or null argument");
  *launch_bytes = LEX_MAX_LAUNCH_BYTES;
  *parser_positions = g->has_parser ? parse_max_positions(g) : 0;
  return 0;
}

// The NUL-separated names of one kind.
static int names_of(const char* grammar, int kind, const uint8_t** names, int* count) {
  const GrammarDesc* g = grammar ? find_grammar(grammar) : nullptr;
  if (!g) return fail(ALPACC_ERR_ARG, "unknown grammar '%s'", grammar ? grammar : "(null)");
  switch (kind) {
    case ALPACC_NAMES_TERMINAL_LEXER: *names = g->terminal_names_lex; *count = g->dead_lex + 1; break;
    case ALPACC_NAMES_TERMINAL: *names = g->terminal_names_both; *count = g->dead_both + 1; break;
    case ALPACC_NAMES_PRODUCTION: *names = g->production_names; *count = g->has_parser ? g->nproductions : 0; break;
    default: return fail(ALPACC_ERR_ARG, "bad name kind %d", kind);
  }
  return 0;
}

int alpacc_hip_name_count(const char* grammar, int kind) {
  const uint8_t* names;
  int count;
  if (int r = names_of(grammar, kind, &names, &count)) return r;
  return count;
}

int alpacc_hip_name(const char* grammar, int kind, uint64_t id, char* buf, size_t cap) {
  const uint8_t* names;
  int count;
  if (!buf) return fail(ALPACC_ERR_ARG, "null argument");
  if (int r = names_of(grammar, kind, &names, &count)) return r;
  if (id >= (uint64_t)count) return fail(ALPACC_ERR_ARG, "id %llu of %d", (unsigned long long)id, count);
  const char* p = reinterpret_cast<const char*>(names);
  for (uint64_t i = 0; i < id; ++i) p += strlen(p) + 1;
  const size_t n = strlen(p);
  if (n + 1 > cap) return fail(ALPACC_ERR_ARG, "name buffer too small");
  memcpy(buf, p, n + 1);
  return (int)n;
}

uint64_t alpacc_hip_grammar_payload_bytes(const char* grammar, int mode, uint64_t n) {
  const GrammarDesc* g = grammar ? find_grammar(grammar) : nullptr;
  if (!g) return 0;
  return mode == ALPACC_MODE_PARSER ? n * (uint64_t)g->terminal_bytes_both : n;
}

uint64_t alpacc_hip_payload_bytes(const alpacc_hip_ctx* h, int mode, uint64_t n) {
  if (!h) return 0;
  return alpacc_hip_grammar_payload_bytes(h->c.g->name, mode, n);
}

void alpacc_hip_free(void* p) { free(p); }

int alpacc_hip_set_timing(alpacc_hip_ctx* h, int on) {
  if (!h) return fail(ALPACC_ERR_ARG, "null argument");
  Context& c = h->c;
  HIP_TRY(hipSetDevice(c.device));
  for (hipEvent_t& ev : c.ev)
    if (!ev) HIP_TRY(hipEventCreate(&ev));
  c.timing = on != 0;
  c.lex_ms = c.parse_ms = 0;
  c.lex_launches = c.parse_launches = 0;
  c.timer.on = on != 0;
  c.timer.reset();
  return 0;
}

int alpacc_hip_timing(const alpacc_hip_ctx* h, const char* stage, double* total_ms, uint64_t* runs) {
  if (!h || !stage || !total_ms || !runs) return fail(ALPACC_ERR_ARG, "null argument");
  const Context& c = h->c;
  if (!strcmp(stage, "lex")) {
    *total_ms = c.lex_ms;
    *runs = c.lex_launches;
  } else if (!strcmp(stage, "parse")) {
    *total_ms = c.parse_ms;
    *runs = c.parse_launches;
  } else if (!c.timer.get(stage, total_ms, runs)) {
    return fail(ALPACC_ERR_ARG, "unknown stage or kernel '%s'", stage);
  }
  return 0;
}

const char* alpacc_hip_timed_kernels(const alpacc_hip_ctx* h) {
  static thread_local std::string out;
  out.clear();
  if (!h) return "";
  for (const std::string& n : h->c.timer.names()) {
    if (!out.empty()) out += ',';
    out += n;
  }
  return out.c_str();
}

const char* const* alpacc_hip_grammars(void) { return grammar_names(); }

}  // extern "C"

namespace alpacc {

hipError_t grant_lds(const void* k, size_t bytes) {
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, size_t> granted;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(mu);
  size_t& g = granted[{dev, k}];
  if (g >= bytes) return hipSuccess;
  e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess) g = bytes;
  return e;
}

hipError_t KernelTimer::begin(hipStream_t st, const char* name) {
  if (!on) return hipSuccess;
  const size_t i = pend_.size();
  while (ev_.size() < 2 * (i + 1)) {
    hipEvent_t e;
    hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) return r;
    ev_.push_back(e);
  }
  pend_.emplace_back(name);
  ended_.push_back(false);
  return hipEventRecord(ev_[2 * i], st);
}

hipError_t KernelTimer::end(hipStream_t st) {
  if (!on || pend_.empty()) return hipSuccess;
  ended_.back() = true;
  return hipEventRecord(ev_[2 * (pend_.size() - 1) + 1], st);
}

hipError_t KernelTimer::collect() {
  for (size_t i = 0; i < pend_.size(); ++i) {
    float ms = 0;
    if (!ended_[i] || hipEventElapsedTime(&ms, ev_[2 * i], ev_[2 * i + 1]) != hipSuccess) continue;
    size_t k = 0;
    while (k < order_.size() && order_[k] != pend_[i]) ++k;
    if (k == order_.size()) {
      order_.push_back(pend_[i]);
      acc_.emplace_back();
    }
    acc_[k].ms += ms;
    acc_[k].runs++;
  }
  pend_.clear();
  ended_.clear();
  return hipSuccess;
}

void KernelTimer::reset() {
  pend_.clear();
  ended_.clear();
  order_.clear();
  acc_.clear();
}

bool KernelTimer::get(const char* name, double* ms, uint64_t* runs) const {
  for (size_t k = 0; k < order_.size(); ++k)
    if (order_[k] == name) {
      *ms = acc_[k].ms;
      *runs = acc_[k].runs;
      return true;
    }
  return false;
}

void KernelTimer::destroy() {
  for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
  ev_.clear();
  reset();
}

}  // namespace alpacc
