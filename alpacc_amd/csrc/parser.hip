// LLP(q,k) parser and parse-tree construction for gfx950.
//
// Semantics are those of the sequential reference (backends/c/parser.c):
//   parse_test      (:49-98)   every position of the framed token string
//                              (start sentinel, tokens, end sentinel) looks
//                              up its q+k window; the stack-change strings
//                              must form one balanced, symbol-matched
//                              bracket sequence; the production strings
//                              concatenate to the leftmost derivation;
//   compute_parents (:103-118) parent of derivation slot i = the last
//                              earlier slot j with e[j] <= e[i], where e is
//                              the exclusive scan of (arity - 1), and the
//                              root is its own parent;
//   compact_tree    (:126-153) terminal slots are dropped and parents
//                              renumbered; tokens get their parent's index.
// Bracket matching uses the same previous-smaller-element search: the left
// bracket matching a right bracket j is the last earlier bracket whose
// depth-before is smaller than j's depth-before.
//
// Every phase is a device-wide pass; scans are reduce-then-scan over tiles
// and the previous-smaller-element queries walk a 32-ary tree of minima.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../../include/alpacc_hip.h"
#include "device.hpp"
#include "parser.hpp"

namespace alpacc {

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return fail(ALPACC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

namespace {

constexpr int PWG = 256;
constexpr int PIPT = 16;
constexpr uint64_t PTILE = (uint64_t)PWG * PIPT;

constexpr uint32_t PARSE_BAD_PAIR = 1u;
constexpr uint32_t PARSE_BAD_DEPTH = 2u;
constexpr uint32_t PARSE_BAD_MATCH = 4u;
constexpr uint32_t PARSE_BAD_TREE = 8u;

struct PairArgs {
  const uint8_t* tok;
  uint64_t m;  // framed length: n + 2
  int q, k, base;
  int empty_t, start_t, end_t;
  const uint16_t* window;
  const int32_t* spans;
};

__device__ __forceinline__ uint32_t window_pair(const PairArgs& a, uint64_t i) {
  uint32_t code = 0;
  const int w = a.q + a.k;
  for (int j = 0; j < w; ++j) {
    int64_t f = (int64_t)i - a.q + j;  // index into the framed string
    uint32_t t;
    if (f < 0 || f >= (int64_t)a.m)
      t = a.empty_t;
    else if (f == 0)
      t = a.start_t;
    else if (f == (int64_t)a.m - 1)
      t = a.end_t;
    else
      t = a.tok[f - 1];
    if (t >= (uint32_t)a.base) return 0xFFFFu;
    code = code * a.base + t;
  }
  return a.window[code];
}

// Block-wide exclusive scan of a u64 pair (lane sums in, block offsets out).
__device__ __forceinline__ void block_scan_pair(uint64_t& x, uint64_t& y, uint64_t& tx, uint64_t& ty) {
  __shared__ uint64_t sx[PWG / WAVE], sy[PWG / WAVE];
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint64_t ix = wave_scan_add64(x), iy = wave_scan_add64(y);
  if (lane == 63) {
    sx[wid] = ix;
    sy[wid] = iy;
  }
  __syncthreads();
  uint64_t ox = 0, oy = 0;
  tx = ty = 0;
  for (int i = 0; i < PWG / WAVE; ++i) {
    if ((uint32_t)i < wid) {
      ox += sx[i];
      oy += sy[i];
    }
    tx += sx[i];
    ty += sy[i];
  }
  x = ox + ix - x;
  y = oy + iy - y;
  __syncthreads();
}

__global__ __launch_bounds__(PWG) void pair_count(PairArgs a, uint64_t* blk, uint32_t* flags) {
  const uint64_t i0 = (uint64_t)blockIdx.x * PTILE + (uint64_t)threadIdx.x * PIPT;
  uint64_t nb = 0, np = 0;
  bool bad = false;
  for (int r = 0; r < PIPT; ++r) {
    uint64_t i = i0 + r;
    if (i < a.m) {
      uint32_t pr = window_pair(a, i);
      if (pr == 0xFFFFu) {
        bad = true;
      } else {
        nb += (uint64_t)a.spans[4 * pr + 1];
        np += (uint64_t)a.spans[4 * pr + 3];
      }
    }
  }
  if (bad) atomicOr(flags, PARSE_BAD_PAIR);
  uint64_t tx, ty;
  block_scan_pair(nb, np, tx, ty);
  if (threadIdx.x == 0) {
    blk[2 * blockIdx.x] = tx;
    blk[2 * blockIdx.x + 1] = ty;
  }
}

// In-place exclusive scan of `cnt` interleaved u64 pairs by one workgroup;
// the totals land after the last pair.
__global__ __launch_bounds__(1024) void scan_pairs_inplace(uint64_t* blk, uint64_t cnt) {
  __shared__ uint64_t wx[16], wy[16];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  uint64_t per = (cnt + 1023) / 1024;
  uint64_t lo = tid * per, hi = lo + per < cnt ? lo + per : cnt;
  uint64_t sx = 0, sy = 0;
  for (uint64_t i = lo; i < hi; ++i) {
    sx += blk[2 * i];
    sy += blk[2 * i + 1];
  }
  uint64_t ix = wave_scan_add64(sx), iy = wave_scan_add64(sy);
  if (lane == 63) {
    wx[wid] = ix;
    wy[wid] = iy;
  }
  __syncthreads();
  uint64_t ox = 0, oy = 0, tx = 0, ty = 0;
  for (int i = 0; i < 16; ++i) {
    if ((uint32_t)i < wid) {
      ox += wx[i];
      oy += wy[i];
    }
    tx += wx[i];
    ty += wy[i];
  }
  uint64_t rx = ox + ix - sx, ry = oy + iy - sy;
  for (uint64_t i = lo; i < hi; ++i) {
    uint64_t vx = blk[2 * i], vy = blk[2 * i + 1];
    blk[2 * i] = rx;
    blk[2 * i + 1] = ry;
    rx += vx;
    ry += vy;
  }
  if (tid == 0) {
    blk[2 * cnt] = tx;
    blk[2 * cnt + 1] = ty;
  }
}

__global__ __launch_bounds__(PWG) void pair_emit(PairArgs a, const uint64_t* blk, const uint16_t* stacks,
                                                  const uint8_t* prods, uint16_t* br, uint8_t* pr) {
  const uint64_t i0 = (uint64_t)blockIdx.x * PTILE + (uint64_t)threadIdx.x * PIPT;
  uint32_t ids[PIPT];
  uint64_t nb = 0, np = 0;
#pragma unroll
  for (int r = 0; r < PIPT; ++r) {
    uint64_t i = i0 + r;
    uint32_t id = 0xFFFFu;
    if (i < a.m) id = window_pair(a, i);
    ids[r] = id;
    if (id != 0xFFFFu) {
      nb += (uint64_t)a.spans[4 * id + 1];
      np += (uint64_t)a.spans[4 * id + 3];
    }
  }
  uint64_t tx, ty;
  block_scan_pair(nb, np, tx, ty);
  uint64_t ob = blk[2 * blockIdx.x] + nb, op = blk[2 * blockIdx.x + 1] + np;
#pragma unroll
  for (int r = 0; r < PIPT; ++r) {
    uint32_t id = ids[r];
    if (id == 0xFFFFu) continue;
    int32_t ss = a.spans[4 * id], sl = a.spans[4 * id + 1];
    int32_t ps = a.spans[4 * id + 2], pl = a.spans[4 * id + 3];
    for (int32_t j = 0; j < sl; ++j) br[ob + j] = stacks[ss + j];
    for (int32_t j = 0; j < pl; ++j) pr[op + j] = prods[ps + j];
    ob += sl;
    op += pl;
  }
}

// ---- generic reduce-then-scan over n items of an int64 value ---------------

template <typename F>
__global__ __launch_bounds__(PWG) void tile_sum(F f, uint64_t n, int64_t* blk) {
  const uint64_t i0 = (uint64_t)blockIdx.x * PTILE + (uint64_t)threadIdx.x * PIPT;
  int64_t s = 0;
  for (int r = 0; r < PIPT; ++r)
    if (i0 + r < n) s += f(i0 + r);
  __shared__ int64_t ws[PWG / WAVE];
  uint64_t inc = wave_scan_add64((uint64_t)s);
  if ((threadIdx.x & 63) == 63) ws[threadIdx.x >> 6] = (int64_t)inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int i = 0; i < PWG / WAVE; ++i) t += ws[i];
    blk[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(1024) void scan_i64_inplace(int64_t* blk, uint64_t cnt) {
  __shared__ int64_t ws[16];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  uint64_t per = (cnt + 1023) / 1024;
  uint64_t lo = tid * per, hi = lo + per < cnt ? lo + per : cnt;
  int64_t s = 0;
  for (uint64_t i = lo; i < hi; ++i) s += blk[i];
  int64_t inc = (int64_t)wave_scan_add64((uint64_t)s);
  if (lane == 63) ws[wid] = inc;
  __syncthreads();
  int64_t o = 0, t = 0;
  for (int i = 0; i < 16; ++i) {
    if ((uint32_t)i < wid) o += ws[i];
    t += ws[i];
  }
  int64_t r = o + inc - s;
  for (uint64_t i = lo; i < hi; ++i) {
    int64_t v = blk[i];
    blk[i] = r;
    r += v;
  }
  if (tid == 0) blk[cnt] = t;
}

template <typename F, typename S>
__global__ __launch_bounds__(PWG) void tile_scan(F f, S sink, uint64_t n, const int64_t* blk) {
  const uint64_t i0 = (uint64_t)blockIdx.x * PTILE + (uint64_t)threadIdx.x * PIPT;
  int64_t v[PIPT];
  int64_t s = 0;
#pragma unroll
  for (int r = 0; r < PIPT; ++r) {
    v[r] = (i0 + r < n) ? f(i0 + r) : 0;
    s += v[r];
  }
  __shared__ int64_t ws[PWG / WAVE];
  int64_t inc = (int64_t)wave_scan_add64((uint64_t)s);
  if ((threadIdx.x & 63) == 63) ws[threadIdx.x >> 6] = inc;
  __syncthreads();
  int64_t o = blk[blockIdx.x];
  for (uint32_t i = 0; i < (threadIdx.x >> 6); ++i) o += ws[i];
  o += inc - s;
#pragma unroll
  for (int r = 0; r < PIPT; ++r) {
    if (i0 + r < n) sink(i0 + r, o, v[r]);
    o += v[r];
  }
}

// ---- 32-ary tree of minima and previous-smaller-element queries -------------

constexpr int MT_MAX_LEVELS = 8;

struct MinTree {
  const int32_t* lv[MT_MAX_LEVELS];
  uint64_t n[MT_MAX_LEVELS];
  int levels;
};

__global__ __launch_bounds__(PWG) void min_level(const int32_t* in, uint64_t n_in, int32_t* out, uint64_t n_out) {
  uint64_t c = (uint64_t)blockIdx.x * PWG + threadIdx.x;
  if (c >= n_out) return;
  uint64_t lo = c * 32, hi = lo + 32 < n_in ? lo + 32 : n_in;
  int32_t m = in[lo];
  for (uint64_t i = lo + 1; i < hi; ++i) m = in[i] < m ? in[i] : m;
  out[c] = m;
}

// Last j < i with a[j] < v (STRICT) or a[j] <= v; -1 when there is none.
template <bool STRICT>
__device__ int64_t prev_smaller(const MinTree& t, uint64_t i, int32_t v) {
  auto ok = [&](int32_t x) { return STRICT ? x < v : x <= v; };
  uint64_t idx = i;
  for (int L = 0; L < t.levels; ++L) {
    const int32_t* a = t.lv[L];
    uint64_t lo = idx & ~31ull;
    for (uint64_t j = idx; j > lo;) {
      --j;
      if (ok(a[j])) {
        uint64_t k = j;
        for (int D = L; D > 0; --D) {
          const int32_t* b = t.lv[D - 1];
          uint64_t c0 = k * 32, c1 = c0 + 32 < t.n[D - 1] ? c0 + 32 : t.n[D - 1];
          uint64_t c = c1;
          while (c > c0) {
            --c;
            if (ok(b[c])) break;
          }
          k = c;
        }
        return (int64_t)k;
      }
    }
    idx >>= 5;
  }
  return -1;
}

// ---- phase functors -----------------------------------------------------------

struct BracketStep {
  const uint16_t* br;
  __device__ int64_t operator()(uint64_t j) const { return (br[j] & 0x8000u) ? 1 : -1; }
};
struct DepthSink {
  int32_t* depth;
  uint32_t* flags;
  __device__ void operator()(uint64_t j, int64_t before, int64_t step) const {
    depth[j] = (int32_t)before;
    if (before + step < 0) atomicOr(flags, PARSE_BAD_DEPTH);
  }
};

__global__ __launch_bounds__(PWG) void match_brackets(const uint16_t* br, MinTree t, uint64_t nb, uint32_t* flags) {
  uint64_t j = (uint64_t)blockIdx.x * PWG + threadIdx.x;
  if (j >= nb) return;
  uint16_t b = br[j];
  if (b & 0x8000u) return;
  int64_t i = prev_smaller<true>(t, j, t.lv[0][j]);
  if (i < 0 || br[i] != (uint16_t)(b | 0x8000u)) atomicOr(flags, PARSE_BAD_MATCH);
}

struct ArityStep {
  const uint8_t* pr;
  const uint8_t* arity;
  __device__ int64_t operator()(uint64_t i) const { return (int64_t)arity[pr[i]] - 1; }
};
struct StoreI32 {
  int32_t* out;
  __device__ void operator()(uint64_t i, int64_t before, int64_t) const { out[i] = (int32_t)before; }
};

__global__ __launch_bounds__(PWG) void parents_kernel(MinTree t, uint64_t np, int32_t* par, uint32_t* flags) {
  uint64_t i = (uint64_t)blockIdx.x * PWG + threadIdx.x;
  if (i >= np) return;
  if (i == 0) {
    par[0] = 0;
    return;
  }
  int64_t j = prev_smaller<false>(t, i, t.lv[0][i]);
  if (j < 0) {
    atomicOr(flags, PARSE_BAD_TREE);
    j = 0;
  }
  par[i] = (int32_t)j;
}

struct TermStep {
  const uint8_t* pr;
  const uint8_t* isterm;
  __device__ int64_t operator()(uint64_t i) const { return isterm[pr[i]]; }
};
struct StoreIncl {
  uint32_t* out;
  __device__ void operator()(uint64_t i, int64_t before, int64_t v) const { out[i] = (uint32_t)(before + v); }
};

__global__ __launch_bounds__(PWG) void compact_kernel(const uint8_t* pr, const uint8_t* isterm, const int32_t* par,
                                                      const uint32_t* tincl, uint64_t np, uint8_t* tree_prod,
                                                      int64_t* tree_par, int64_t* tok_par) {
  uint64_t i = (uint64_t)blockIdx.x * PWG + threadIdx.x;
  if (i >= np) return;
  uint8_t p = pr[i];
  int32_t po = par[i];
  int64_t pn = (int64_t)po - (int64_t)tincl[po];
  uint32_t ti = tincl[i];
  if (isterm[p]) {
    tok_par[ti - 1] = pn;
  } else {
    tree_prod[i - ti] = p;
    tree_par[i - ti] = pn;
  }
}

inline uint32_t blocks_for(uint64_t n, uint64_t per) { return (uint32_t)((n + per - 1) / per); }

template <typename F, typename S>
int device_scan(F f, S s, uint64_t n, int64_t* blk, hipStream_t st, int64_t* total_host) {
  uint32_t nb = blocks_for(n, PTILE);
  hipLaunchKernelGGL(tile_sum<F>, dim3(nb), dim3(PWG), 0, st, f, n, blk);
  hipLaunchKernelGGL(scan_i64_inplace, dim3(1), dim3(1024), 0, st, blk, (uint64_t)nb);
  hipLaunchKernelGGL((tile_scan<F, S>), dim3(nb), dim3(PWG), 0, st, f, s, n, blk);
  HIP_TRY(hipGetLastError());
  if (total_host) {
    HIP_TRY(hipMemcpyAsync(total_host, blk + nb, 8, hipMemcpyDeviceToHost, st));
  }
  return 0;
}

// Builds the levels above lv0 inside `buf` (grown as needed).
int build_min_tree(const int32_t* lv0, uint64_t n, DevBuf& buf, hipStream_t st, MinTree* t) {
  uint64_t total = 0, m = n;
  int levels = 1;
  while (m > 32 && levels < MT_MAX_LEVELS) {
    m = (m + 31) / 32;
    total += m;
    ++levels;
  }
  if (m > 32) return fail(ALPACC_ERR_LIMIT, "min tree too deep");
  int r;
  if ((r = buf.reserve(total * 4 + 16))) return r;
  t->levels = levels;
  t->lv[0] = lv0;
  t->n[0] = n;
  int32_t* p = buf.as<int32_t>();
  for (int L = 1; L < levels; ++L) {
    uint64_t nn = (t->n[L - 1] + 31) / 32;
    hipLaunchKernelGGL(min_level, dim3(blocks_for(nn, PWG)), dim3(PWG), 0, st, t->lv[L - 1], t->n[L - 1], p, nn);
    t->lv[L] = p;
    t->n[L] = nn;
    p += nn;
  }
  HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace

int parse_on_device(Context& c, const uint8_t* d_tok, uint64_t n, bool tree, hipStream_t st, uint64_t* num_prods) {
  const GrammarDesc* g = c.g;
  *num_prods = 0;
  const uint64_t m = n + 2;
  int r;
  PairArgs a{d_tok, m, g->q, g->k, g->window_base, g->empty_t, g->start_t, g->end_t, c.d_window, c.d_spans};
  uint32_t* flags = reinterpret_cast<uint32_t*>(c.small.as<uint64_t>() + 2);
  if ((r = zero_words(flags, 4, st))) return r;
  const uint32_t nblk = blocks_for(m, PTILE);
  if ((r = c.ps_a.reserve((size_t)(nblk + 1) * 16))) return r;
  uint64_t* blk = c.ps_a.as<uint64_t>();
  if (c.timing) HIP_TRY(hipEventRecord(c.ev[2], st));
  hipLaunchKernelGGL(pair_count, dim3(nblk), dim3(PWG), 0, st, a, blk, flags);
  hipLaunchKernelGGL(scan_pairs_inplace, dim3(1), dim3(1024), 0, st, blk, (uint64_t)nblk);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(c.host_small, blk + 2 * nblk, 16, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(c.host_small + 2, flags, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  const uint64_t nbr = c.host_small[0], npr = c.host_small[1];
  if (reinterpret_cast<uint32_t*>(c.host_small + 2)[0]) return ALPACC_REJECT;
  if (nbr > 0x7FFFFFFFull || npr > 0x7FFFFFFFull)
    return fail(ALPACC_ERR_LIMIT, "derivation of %llu brackets / %llu productions exceeds 2^31",
                (unsigned long long)nbr, (unsigned long long)npr);
  if ((r = c.ps_b.reserve(npr + 16)) || (r = c.ps_c.reserve(nbr * 2 + 16))) return r;
  uint8_t* pr = c.ps_b.as<uint8_t>();
  uint16_t* br = c.ps_c.as<uint16_t>();
  hipLaunchKernelGGL(pair_emit, dim3(nblk), dim3(PWG), 0, st, a, blk, c.d_stacks, c.d_productions, br, pr);
  HIP_TRY(hipGetLastError());

  // bracket depths and matching
  if (nbr) {
    if ((r = c.ps_d.reserve(nbr * 4 + 16)) || (r = c.ps_i.reserve((size_t)(blocks_for(nbr, PTILE) + 1) * 8 + 16)))
      return r;
    int32_t* depth = c.ps_d.as<int32_t>();
    int64_t* sblk = c.ps_i.as<int64_t>();
    if ((r = device_scan(BracketStep{br}, DepthSink{depth, flags}, nbr, sblk, st, (int64_t*)(c.host_small + 3))))
      return r;
    MinTree t{};
    if ((r = build_min_tree(depth, nbr, c.ps_e, st, &t))) return r;
    hipLaunchKernelGGL(match_brackets, dim3(blocks_for(nbr, PWG)), dim3(PWG), 0, st, br, t, nbr, flags);
    HIP_TRY(hipGetLastError());
  } else {
    c.host_small[3] = 0;
  }
  HIP_TRY(hipMemcpyAsync(c.host_small + 2, flags, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (reinterpret_cast<uint32_t*>(c.host_small + 2)[0] || (int64_t)c.host_small[3] != 0) return ALPACC_REJECT;
  *num_prods = npr;
  if (!tree) return ALPACC_VALID;

  // parents: previous smaller-or-equal element of exscan(arity - 1)
  if ((r = c.ps_f.reserve(npr * 4 + 16)) || (r = c.ps_g.reserve(npr * 4 + 16)) ||
      (r = c.ps_i.reserve((size_t)(blocks_for(npr, PTILE) + 1) * 8 + 16)))
    return r;
  int32_t* e = c.ps_f.as<int32_t>();
  int64_t* sblk = c.ps_i.as<int64_t>();
  if ((r = device_scan(ArityStep{pr, c.d_arity}, StoreI32{e}, npr, sblk, st, nullptr))) return r;
  MinTree t{};
  if ((r = build_min_tree(e, npr, c.ps_e, st, &t))) return r;
  if ((r = c.ps_h.reserve(npr * 4 + 16))) return r;
  int32_t* par = c.ps_h.as<int32_t>();
  hipLaunchKernelGGL(parents_kernel, dim3(blocks_for(npr, PWG)), dim3(PWG), 0, st, t, npr, par, flags);
  HIP_TRY(hipGetLastError());
  // compaction
  uint32_t* tincl = c.ps_g.as<uint32_t>();
  if ((r = device_scan(TermStep{pr, c.d_isterm}, StoreIncl{tincl}, npr, sblk, st, (int64_t*)(c.host_small + 3))))
    return r;
  HIP_TRY(hipStreamSynchronize(st));
  if ((uint64_t)c.host_small[3] != n) return ALPACC_REJECT;
  const uint64_t nodes = npr - n;
  if ((r = c.tr_prod.reserve(nodes + 16)) || (r = c.tr_parent.reserve(nodes * 8 + 16)) ||
      (r = c.tr_tokpar.reserve(n * 8 + 16)))
    return r;
  hipLaunchKernelGGL(compact_kernel, dim3(blocks_for(npr, PWG)), dim3(PWG), 0, st, pr, c.d_isterm, par, tincl, npr,
                     c.tr_prod.as<uint8_t>(), c.tr_parent.as<int64_t>(), c.tr_tokpar.as<int64_t>());
  HIP_TRY(hipGetLastError());
  if (c.timing) HIP_TRY(hipEventRecord(c.ev[3], st));
  HIP_TRY(hipMemcpyAsync(c.host_small + 2, flags, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (c.timing) {
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, c.ev[2], c.ev[3]));
    c.parse_ms += ms;
    c.parse_launches++;
  }
  if (reinterpret_cast<uint32_t*>(c.host_small + 2)[0]) return ALPACC_REJECT;
  return ALPACC_VALID;
}

}  // namespace alpacc
