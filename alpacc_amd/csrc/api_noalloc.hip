// alpacc_hip_lex_device_noalloc: the device-memory lexer entry point
// without any allocation, for callers that cannot allocate (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Its scratch memory -- the lexer's lookback words, tile and lane prefixes,
// the launch metadata and the workgroup ticket -- is one file-scope device
// array, `scratch`, of NOALLOC_SCRATCH bytes on every device the library
// runs on (held until the process ends); inputs and outputs are the
// caller's.  Host code never passes the array's name to a kernel: the
// address of each device's copy is looked up once per device
// (hipGetSymbolAddress with that device current) and kept, and that pointer
// is what the kernels receive.  Every call on a device shares its array, so
// calls that overlap on one device race; each call clears what must start
// at zero (the lookback words, the ticket, the verdict flags) with a fill
// kernel first.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/alpacc_hip.h"
#include "lexer.hpp"
#include "runtime.hpp"

namespace alpacc {

namespace {

constexpr size_t NOALLOC_SCRATCH = ALPACC_NOALLOC_SCRATCH_BYTES;
constexpr uint64_t NOALLOC_MAX_INPUT = ALPACC_NOALLOC_MAX_INPUT;

__device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
uint8_t* g_scratch[MAX_DEVICES] = {};

// The address of the current device's copy of `scratch` (looked up on the
// first call on the device).
int scratch_of(int dev, uint8_t** out) {
  if (dev < 0 || dev >= MAX_DEVICES) return fail(ALPACC_ERR_ARG, "device %d beyond %d", dev, MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(ALPACC_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch[dev] = static_cast<uint8_t*>(p);
  }
  *out = g_scratch[dev];
  return 0;
}

// The launch metadata of one input (what the batch path uploads from its
// pinned host mirror), written on the device.
__global__ void noalloc_meta(uint8_t* meta, BatchLayout L, uint64_t n, uint32_t nchunks, uint32_t ntiles) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint64_t* seg = reinterpret_cast<uint64_t*>(meta + L.off_seg);
  uint32_t* cf = reinterpret_cast<uint32_t*>(meta + L.off_cf);
  seg[0] = 0;
  seg[1] = n;
  cf[0] = 0;
  cf[1] = nchunks;
  *reinterpret_cast<uint32_t*>(meta + L.off_nt) = ntiles;
  *reinterpret_cast<uint32_t*>(meta + L.off_live) = 1;
}

}  // namespace

}  // namespace alpacc

using namespace alpacc;

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return fail(ALPACC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

extern "C" int alpacc_hip_lex_device_noalloc(alpacc_hip_ctx* h, int mode, const uint8_t* d_str, uint64_t n,
                                             void* d_terminals, int64_t* d_starts, void* d_lengths, uint64_t cap,
                                             uint64_t* count, void* stream) {
  if (!h || !count || (n && (!d_str || !d_terminals || !d_starts || !d_lengths)))
    return fail(ALPACC_ERR_ARG, "null argument");
  if (mode != ALPACC_MODE_LEXER && mode != ALPACC_MODE_BOTH) return fail(ALPACC_ERR_ARG, "bad lexer mode %d", mode);
  *count = 0;
  if (n > NOALLOC_MAX_INPUT)
    return fail(ALPACC_ERR_LIMIT, "input of %llu bytes beyond the %llu bytes of the allocation-free lexer",
                (unsigned long long)n, (unsigned long long)NOALLOC_MAX_INPUT);
  if (cap < n) return fail(ALPACC_ERR_ARG, "output capacity %llu below the input's %llu bytes", (unsigned long long)cap,
                           (unsigned long long)n);
  if (n == 0) return ALPACC_VALID;
  Context& c = h->c;
  const GrammarDesc* g = c.g;
  if (g->nendo > 0x8000) return fail(ALPACC_ERR_LIMIT, "%d lexer endomorphisms exceed 32768", g->nendo);
  HIP_TRY(hipSetDevice(c.device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c.stream;
  uint8_t* sc = nullptr;
  if (int r = scratch_of(c.device, &sc)) return r;

  // scratch layout: ticket (16 B) | metadata (BatchLayout) | lookback words
  // u64[ntiles] | tile_endo u32[ntiles] | tile_prefix u16[ntiles] |
  // tile_blk u32[2 * scan blocks] | lane_prefix u16[ntiles * LEX_WG]
  const uint64_t ipt = (uint64_t)c.lex_ipt;
  const uint64_t nchunks = (n + ipt - 1) / ipt;
  const uint64_t ntiles = (nchunks + LEX_WG - 1) / LEX_WG;
  const BatchLayout L(1, ntiles);
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t off_meta = 256, off_st = al(off_meta + L.total);
  const size_t nblk = ((size_t)ntiles + 255) / 256;
  const size_t te_off = off_st + ntiles * 8, tp_off = te_off + ntiles * 4, tb_off = al(tp_off + ntiles * 2),
               lp_off = al(tb_off + nblk * 8), end = lp_off + ntiles * LEX_WG * 2;
  if (end > NOALLOC_SCRATCH) return fail(ALPACC_ERR_LIMIT, "scratch of %zu bytes needed", end);
  uint8_t* dm = sc + off_meta;

  LexParams p{};
  p.in = d_str;
  p.in_bytes = n;
  p.seg_off = reinterpret_cast<const uint64_t*>(dm + L.off_seg);
  p.chunk_first = reinterpret_cast<const uint32_t*>(dm + L.off_cf);
  p.nt_dev = reinterpret_cast<const uint32_t*>(dm + L.off_nt);
  p.tile_seg = reinterpret_cast<uint32_t*>(dm + L.off_tseg);
  p.nseg = 1;
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
  const int lb = g->length_bytes;
  p.max_len = lb >= 8 ? ~0ull : ((1ull << (8 * lb)) - 1);
  p.length_bytes = lb;
  p.terminals = static_cast<uint8_t*>(d_terminals);
  p.starts = d_starts;
  p.lengths = d_lengths;
  p.seg_flags = reinterpret_cast<uint32_t*>(dm + L.off_flags);
  p.seg_tok = reinterpret_cast<uint64_t*>(dm + L.off_tok);
  p.ticket = reinterpret_cast<uint32_t*>(sc);
  p.count_status = reinterpret_cast<uint64_t*>(sc + off_st);
  p.tile_endo = reinterpret_cast<uint32_t*>(sc + te_off);
  p.tile_prefix = reinterpret_cast<uint16_t*>(sc + tp_off);
  p.tile_blk = reinterpret_cast<uint32_t*>(sc + tb_off);
  p.lane_prefix = reinterpret_cast<uint16_t*>(sc + lp_off);
  p.ntiles = (uint32_t)ntiles;
  p.ipt = c.lex_ipt;
  p.lb_win = c.lex_lb;
  p.is_final = 1;
  // persistent grids of the context's shape (computed once per shape; no
  // allocation)
  if (!c.lex_grid) HIP_TRY(lex_grid_size(p, c.lex_wpc, &c.lex_grid, &c.lex_grid2));

  hipLaunchKernelGGL(noalloc_meta, dim3(1), dim3(64), 0, st, dm, L, n, (uint32_t)nchunks, (uint32_t)ntiles);
  HIP_TRY(hipGetLastError());
  int r;
  if ((r = zero_ranges({{sc, 16}, {p.count_status, ntiles * 8}, {p.seg_flags, 4}}, st)))
    return r;
  KernelTimer off;  // no events: they would be created on first use
  HIP_TRY(lex_launch(p, c.lex_grid, c.lex_grid2, st, off));
  // the verdict and the token count, into host memory of the caller's stack
  uint64_t tok[2] = {0, 0};
  uint32_t flags = 0;
  HIP_TRY(hipMemcpyAsync(tok, p.seg_tok, 16, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(&flags, p.seg_flags, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  *count = tok[1] - tok[0];
  return flags ? ALPACC_REJECT : ALPACC_VALID;
}
