// run_case: one request payload -> one wire response record.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/alpacc_hip.h"
#include "runtime.hpp"

namespace alpacc {
int lexer_record(Context& c, uint64_t n, const uint8_t* payload, std::vector<uint8_t>& rec);
int parser_record(Context& c, uint64_t n, const uint8_t* payload, std::vector<uint8_t>& rec);
int both_record(Context& c, uint64_t n, const uint8_t* payload, std::vector<uint8_t>& rec);
}  // namespace alpacc

using namespace alpacc;

extern "C" int alpacc_hip_run_case(alpacc_hip_ctx* h, int mode, uint64_t n, const uint8_t* payload, uint8_t** record,
                                   uint64_t* record_len) {
  if (!h || !record || !record_len || (n && !payload)) return fail(ALPACC_ERR_ARG, "null argument");
  *record = nullptr;
  *record_len = 0;
  Context& c = h->c;
  hipError_t e = hipSetDevice(c.device);
  if (e != hipSuccess) return fail(ALPACC_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
  std::vector<uint8_t> rec;
  int r;
  switch (mode) {
    case ALPACC_MODE_LEXER: r = lexer_record(c, n, payload, rec); break;
    case ALPACC_MODE_PARSER: r = parser_record(c, n, payload, rec); break;
    case ALPACC_MODE_BOTH: r = both_record(c, n, payload, rec); break;
    default: return fail(ALPACC_ERR_ARG, "bad mode %d", mode);
  }
  if (r < 0) return r;
  uint8_t* out = static_cast<uint8_t*>(malloc(rec.size()));
  if (!out) return fail(ALPACC_ERR_NOMEM, "host allocation failed");
  memcpy(out, rec.data(), rec.size());
  *record = out;
  *record_len = rec.size();
  return r;
}
