// Serving runtime: continuous batching over resident grammars.
//
// Admission: a batch takes queued requests while their estimated device
// memory (64 bytes per input byte) fits both the server's budget (64 GiB by
// default) and the lane: what its grow-only workspaces already hold, plus
// the device's free memory (hipMemGetInfo) less what other lanes' admitted
// batches may still allocate and a 1 GiB reserve.  A request larger than
// that is a batch of its own.
//
// The reference serves one request at a time over the wire protocol
// (backends/c/cli.c:167-189, docs/wire-protocols.md): every request is a
// separate run_case.  Here every resident grammar has `lanes` contexts, each
// with its own stream, device buffers and worker thread.  A worker that is
// free takes every request queued for its grammar in one mode -- up to the
// device-memory budget still unadmitted -- and answers them with one batched
// launch sequence (batch_records).  Requests that arrive while other batches
// are on the GPU wait only for the next free lane, and the batches of
// different lanes overlap on the device.  Every record is the one
// alpacc_hip_run_case returns for the request alone.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/alpacc_hip.h"
#include "runtime.hpp"

using Clock = std::chrono::steady_clock;

namespace {

struct Req {
  uint32_t g = 0;
  int mode = 0;
  uint64_t seq = 0;  // arrival order
  const uint8_t* data = nullptr;
  uint64_t n = 0;
  std::vector<uint8_t> own;  // payload copy of a submitted request
  uint64_t est = 0;          // device bytes admitted for it
  std::shared_ptr<uint8_t> buf;  // its batch's records; this one is [rec_at, rec_at + rec_len)
  size_t rec_at = 0, rec_len = 0;
  int status = 0;  // 0 waiting, 1 answered, < 0 error
  std::string err;
  Clock::time_point done;
};

// Device bytes a request may hold while its batch runs: the input, up to one
// token per byte (terminal, start, length, token parent: 19 bytes), the
// tree and the parser's intermediates; a conservative 64 per input byte.
uint64_t estimate(uint64_t n) { return n * 64 + 4096; }

// Payload bytes a batch takes at most (one larger request is a batch of its
// own): past a few MiB a batch no longer gains throughput, only latency and
// workspace growth.
constexpr uint64_t kBatchBytes = 8ull << 20;

// Device memory left free for everything else (other processes, the HIP
// runtime) when admission looks at the device's free memory.
constexpr uint64_t kReserve = 1ull << 30;

// Device bytes a lane's context already holds in its grow-only workspaces:
// a batch within them allocates nothing.
uint64_t lane_capacity(alpacc_hip_ctx* h) {
  uint64_t cap = 0;
  for (alpacc::DevBuf* b : h->c.all_buffers()) cap += b->cap;
  return cap;
}

}  // namespace

// Record buffers kept for reuse.  A batch's records take tens of MB; a
// fresh block of that size is mapped from the system and unmapped when
// freed, and the page faults of its first writes and the unmapping's TLB
// shootdowns stall every lane's thread (a server's p99 went from ~20 ms to
// ~2 ms at 150k req/s with large blocks kept).  So blocks go back to this
// pool (up to kKeep of them) when their batch's last request is collected,
// and a batch takes the first idle block large enough.  Only large blocks
// (kMin and up; smaller ones come from the heap and are cheap) are pooled.
// Each block starts with a 16-byte header holding its capacity (0: not
// pooled).
struct RecPool {
  static constexpr size_t kKeep = 64, kMin = 8ull << 20;
  std::mutex mu;
  std::vector<uint8_t*> idle;  // block bases
  ~RecPool() {
    for (uint8_t* b : idle) free(b);
  }
  static size_t cap_of(const uint8_t* base) {
    size_t c;
    memcpy(&c, base, sizeof c);
    return c;
  }
  uint8_t* get(size_t n) {
    if (n < kMin) {
      uint8_t* b = static_cast<uint8_t*>(malloc(n + 16));
      if (!b) return nullptr;
      const size_t none = 0;
      memcpy(b, &none, sizeof none);
      return b + 16;
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      for (size_t i = 0; i < idle.size(); ++i)
        if (cap_of(idle[i]) >= n) {
          uint8_t* b = idle[i];
          idle.erase(idle.begin() + (ptrdiff_t)i);
          return b + 16;
        }
    }
    const size_t cap = n + n / 8;  // room for somewhat larger batches later
    uint8_t* b = static_cast<uint8_t*>(malloc(cap + 16));
    if (!b) return nullptr;
    memcpy(b, &cap, sizeof cap);
    return b + 16;
  }
  void put(uint8_t* p) {
    if (!p) return;
    uint8_t* b = p - 16;
    if (cap_of(b) == 0) {
      free(b);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      if (idle.size() < kKeep) {
        idle.push_back(b);
        return;
      }
    }
    free(b);
  }
};

struct alpacc_hip_server {
  std::vector<std::string> names;
  int device = 0;
  int lanes = 1;
  uint64_t budget = 0, max_batch = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::vector<std::deque<Req*>> q;  // [grammar * 3 + mode]
  uint64_t in_use = 0, seq = 0;
  uint64_t growing = 0;  // device bytes that admitted batches may still allocate
  bool stop = false;
  std::vector<std::thread> workers;
  std::vector<alpacc_hip_ctx*> ctxs;
  std::unordered_map<uint64_t, std::unique_ptr<Req>> tickets;
  uint64_t next_ticket = 1;
  uint64_t n_batches = 0, n_reqs = 0, biggest = 0, peak_in_use = 0;
  std::atomic<bool> graphs{false};  // combined batches through each lane's captured graph
  // record blocks for reuse; requests hold their batch's block (a shared
  // owner of the pool too, so blocks released after the server is gone
  // still find it)
  std::shared_ptr<RecPool> pool = std::make_shared<RecPool>();
};

namespace {

// Queue of grammar g to serve next: the one whose head arrived first among
// those whose head fits the budget (or anything when nothing is admitted).
int pick(alpacc_hip_server* s, uint32_t g) {
  int best = -1;
  for (int m = 0; m < 3; ++m) {
    const auto& d = s->q[g * 3 + m];
    if (d.empty()) continue;
    if (s->in_use && s->in_use + d.front()->est > s->budget) continue;
    if (best < 0 || d.front()->seq < s->q[g * 3 + best].front()->seq) best = m;
  }
  return best;
}

void worker(alpacc_hip_server* s, uint32_t g, alpacc_hip_ctx* ctx) {
  (void)hipSetDevice(s->device);
  std::vector<Req*> batch;
  // the batch's payloads, packed in pinned memory (the upload is then a
  // direct DMA instead of the runtime's staged copy from pageable memory);
  // reused batch after batch, each upload done before batch_records returns
  alpacc::PinnedBuf packed;
  // pinned buffers sized for a full batch before the first request: growing
  // one (hipHostMalloc pins its pages) stalls for milliseconds, so it
  // happens only for a request beyond the batch cap, not while serving
  // ordinary ones (the outputs read back take about four bytes per input
  // byte on JSON)
  (void)packed.reserve(kBatchBytes + 16);
  (void)ctx->c.h_out.reserve(4 * kBatchBytes);
  std::vector<uint64_t> off;
  std::vector<size_t> rec_off;
  for (;;) {
    int mode;
    uint64_t est = 0, grow = 0;
    // what the lane holds already; beyond it a batch needs free device memory
    const uint64_t held = lane_capacity(ctx);
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return s->stop || pick(s, g) >= 0; });
      mode = pick(s, g);
      if (mode < 0) {  // stopping with nothing left for this grammar
        packed.release();
        return;
      }
      auto& d = s->q[g * 3 + mode];
      batch.clear();
      uint64_t bytes = 0, room = held;
      bool asked = false;
      while (!d.empty() && (s->max_batch == 0 || batch.size() < s->max_batch)) {
        Req* r = d.front();
        if (!batch.empty() && (s->in_use + est + r->est > s->budget || bytes + r->n > kBatchBytes)) break;
        if (!batch.empty() && est + r->est > room) {
          // admission by free HBM: the device's free memory, less what
          // other lanes' admitted batches may still allocate, less a reserve
          if (!asked) {
            size_t free_b = 0, total_b = 0;
            if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
              const uint64_t avail = free_b > s->growing + kReserve ? free_b - s->growing - kReserve : 0;
              room = held + avail;
            }
            asked = true;
          }
          if (est + r->est > room) break;
        }
        batch.push_back(r);
        est += r->est;
        bytes += r->n;
        d.pop_front();
      }
      grow = est > held ? est - held : 0;
      s->in_use += est;
      s->growing += grow;
      if (s->in_use > s->peak_in_use) s->peak_in_use = s->in_use;
    }
    off.assign(1, 0);
    uint64_t total = 0;
    for (Req* r : batch) total += r->n;
    int prc = packed.reserve(total + 16);
    for (Req* r : batch) {
      if (prc == 0 && r->n) memcpy(packed.as<uint8_t>() + off.back(), r->data, r->n);
      off.push_back(off.back() + r->n);
    }
    uint8_t* out = nullptr;
    size_t out_len = 0;
    ctx->c.use_graphs = s->graphs.load();
    const int rc = prc < 0 ? prc
                           : alpacc::batch_records(
                                 ctx->c, mode, packed.as<uint8_t>(), off.data(), batch.size(), &out, &out_len, &rec_off,
                                 [](void* pool, size_t n) { return static_cast<RecPool*>(pool)->get(n); },
                                 s->pool.get());
    const std::string err = rc < 0 ? alpacc_hip_last_error() : "";
    std::shared_ptr<RecPool> pool = s->pool;
    std::shared_ptr<uint8_t> shared(out, [pool](uint8_t* p) { pool->put(p); });
    const Clock::time_point now = Clock::now();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      for (size_t i = 0; i < batch.size(); ++i) {
        Req* r = batch[i];
        if (rc < 0) {
          r->status = rc;
          r->err = err;
        } else {
          const size_t a = rec_off[i], b = i + 1 < batch.size() ? rec_off[i + 1] : out_len;
          r->buf = shared;
          r->rec_at = a;
          r->rec_len = b - a;
          r->status = 1;
        }
        r->done = now;
      }
      s->in_use -= est;
      s->growing -= grow;
      s->n_batches++;
      s->n_reqs += batch.size();
      if (batch.size() > s->biggest) s->biggest = batch.size();
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
}

void enqueue(alpacc_hip_server* s, Req* r) {
  r->est = estimate(r->n);
  r->seq = s->seq++;
  s->q[r->g * 3 + r->mode].push_back(r);
}

}  // namespace

extern "C" {

int alpacc_hip_server_open(const char* const* grammars, int ngrammars, int device, int lanes, uint64_t hbm_budget,
                           uint64_t max_batch, alpacc_hip_server** out) {
  return alpacc_hip_server_open_variant(grammars, ngrammars, device, lanes, hbm_budget, max_batch, 0, 0, out);
}

int alpacc_hip_server_open_variant(const char* const* grammars, int ngrammars, int device, int lanes,
                                   uint64_t hbm_budget, uint64_t max_batch, int index32, int length_bits,
                                   alpacc_hip_server** out) {
  if (!grammars || ngrammars <= 0 || lanes <= 0 || !out) return alpacc::fail(ALPACC_ERR_ARG, "bad server arguments");
  *out = nullptr;
  auto* s = new alpacc_hip_server();
  s->device = device;
  s->lanes = lanes;
  s->budget = hbm_budget ? hbm_budget : (64ull << 30);
  s->max_batch = max_batch;
  s->q.resize((size_t)ngrammars * 3);
  for (int g = 0; g < ngrammars; ++g) {
    s->names.emplace_back(grammars[g]);
    for (int l = 0; l < lanes; ++l) {
      alpacc_hip_ctx* c = nullptr;
      const int r = alpacc_hip_open_variant(grammars[g], device, index32, length_bits, &c);
      if (r < 0) {
        for (alpacc_hip_ctx* x : s->ctxs) alpacc_hip_close(x);
        delete s;
        return r;
      }
      s->ctxs.push_back(c);
    }
  }
  for (int g = 0; g < ngrammars; ++g)
    for (int l = 0; l < lanes; ++l) s->workers.emplace_back(worker, s, (uint32_t)g, s->ctxs[(size_t)g * lanes + l]);
  *out = s;
  return 0;
}

void alpacc_hip_server_close(alpacc_hip_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (std::thread& t : s->workers) t.join();
  for (alpacc_hip_ctx* c : s->ctxs) alpacc_hip_close(c);
  delete s;
}

int alpacc_hip_server_submit(alpacc_hip_server* s, int grammar, int mode, const uint8_t* payload, uint64_t n,
                             uint64_t* ticket) {
  if (!s || !ticket || (n && !payload)) return alpacc::fail(ALPACC_ERR_ARG, "null argument");
  if (grammar < 0 || (size_t)grammar >= s->names.size()) return alpacc::fail(ALPACC_ERR_ARG, "bad grammar %d", grammar);
  if (mode != ALPACC_MODE_LEXER && mode != ALPACC_MODE_PARSER && mode != ALPACC_MODE_BOTH)
    return alpacc::fail(ALPACC_ERR_ARG, "bad mode %d", mode);
  auto r = std::make_unique<Req>();
  r->g = (uint32_t)grammar;
  r->mode = mode;
  r->own.assign(payload, payload + n);
  r->data = r->own.data();
  r->n = n;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    *ticket = s->next_ticket++;
    enqueue(s, r.get());
    s->tickets.emplace(*ticket, std::move(r));
  }
  s->cv_work.notify_all();
  return 0;
}

int alpacc_hip_server_wait(alpacc_hip_server* s, uint64_t ticket, uint8_t** record, uint64_t* record_len) {
  if (!s || !record || !record_len) return alpacc::fail(ALPACC_ERR_ARG, "null argument");
  std::unique_ptr<Req> r;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return alpacc::fail(ALPACC_ERR_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    Req* p = it->second.get();
    s->cv_done.wait(lk, [&] { return p->status != 0; });
    r = std::move(it->second);
    s->tickets.erase(it);
  }
  if (r->status < 0) return alpacc::fail(r->status, "%s", r->err.c_str());
  uint8_t* o = static_cast<uint8_t*>(malloc(r->rec_len ? r->rec_len : 1));
  if (!o) return alpacc::fail(ALPACC_ERR_NOMEM, "host allocation failed");
  if (r->rec_len) memcpy(o, r->buf.get() + r->rec_at, r->rec_len);
  *record = o;
  *record_len = r->rec_len;
  return 0;
}

int alpacc_hip_server_replay(alpacc_hip_server* s, uint64_t count, const int32_t* grammar, const int32_t* mode,
                             const uint8_t* packed, const uint64_t* offsets, const uint64_t* arrival_ns,
                             uint64_t* latency_ns, uint8_t** records, uint64_t* records_len) {
  if (!s || !grammar || !mode || !offsets || !arrival_ns || !latency_ns || !records || !records_len)
    return alpacc::fail(ALPACC_ERR_ARG, "null argument");
  for (uint64_t i = 0; i < count; ++i) {
    if (grammar[i] < 0 || (size_t)grammar[i] >= s->names.size() || mode[i] < 0 || mode[i] > 2 ||
        offsets[i + 1] < offsets[i])
      return alpacc::fail(ALPACC_ERR_ARG, "bad request %llu", (unsigned long long)i);
  }
  std::vector<Req> reqs(count);
  for (uint64_t i = 0; i < count; ++i) {
    reqs[i].g = (uint32_t)grammar[i];
    reqs[i].mode = mode[i];
    reqs[i].data = packed + offsets[i];
    reqs[i].n = offsets[i + 1] - offsets[i];
  }
  const Clock::time_point t0 = Clock::now();
  std::thread submitter([&] {
    for (uint64_t i = 0; i < count; ++i) {
      std::this_thread::sleep_until(t0 + std::chrono::nanoseconds(arrival_ns[i]));
      {
        std::lock_guard<std::mutex> lk(s->mu);
        enqueue(s, &reqs[i]);
      }
      s->cv_work.notify_all();
    }
  });
  {
    std::unique_lock<std::mutex> lk(s->mu);
    for (uint64_t i = 0; i < count; ++i) s->cv_done.wait(lk, [&] { return reqs[i].status != 0; });
  }
  submitter.join();
  size_t total = 0;
  for (uint64_t i = 0; i < count; ++i) {
    if (reqs[i].status < 0) return alpacc::fail(reqs[i].status, "%s", reqs[i].err.c_str());
    total += reqs[i].rec_len;
    latency_ns[i] = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                        reqs[i].done - (t0 + std::chrono::nanoseconds(arrival_ns[i])))
                        .count();
  }
  uint8_t* o = static_cast<uint8_t*>(malloc(total ? total : 1));
  if (!o) return alpacc::fail(ALPACC_ERR_NOMEM, "host allocation failed");
  size_t at = 0;
  for (uint64_t i = 0; i < count; ++i) {
    if (reqs[i].rec_len) memcpy(o + at, reqs[i].buf.get() + reqs[i].rec_at, reqs[i].rec_len);
    at += reqs[i].rec_len;
  }
  *records = o;
  *records_len = total;
  return 0;
}

int alpacc_hip_server_set_graphs(alpacc_hip_server* s, int on) {
  if (!s) return alpacc::fail(ALPACC_ERR_ARG, "null argument");
  s->graphs.store(on != 0);
  return 0;
}

int alpacc_hip_server_stats(alpacc_hip_server* s, uint64_t* batches, uint64_t* requests, uint64_t* max_batch) {
  if (!s || !batches || !requests || !max_batch) return alpacc::fail(ALPACC_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *batches = s->n_batches;
  *requests = s->n_reqs;
  *max_batch = s->biggest;
  return 0;
}

}  // extern "C"
