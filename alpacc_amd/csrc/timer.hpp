// Per-kernel durations from HIP events recorded on the launch stream.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace alpacc {

// While `on`, begin()/end() bracket each launch with a pair of events on its
// stream; collect() (after the caller has synchronised that stream) adds the
// elapsed times of the completed intervals to per-name totals and drops the
// rest (a launch abandoned on an error path).  Off, every call is a no-op.
class KernelTimer {
 public:
  bool on = false;
  hipError_t begin(hipStream_t st, const char* name);
  hipError_t end(hipStream_t st);
  hipError_t collect();
  void reset();
  bool get(const char* name, double* ms, uint64_t* runs) const;
  const std::vector<std::string>& names() const { return order_; }
  void destroy();

 private:
  struct Acc {
    double ms = 0;
    uint64_t runs = 0;
  };
  std::vector<hipEvent_t> ev_;     // pool: interval i uses ev_[2i], ev_[2i+1]
  std::vector<std::string> pend_;  // names of the recorded, uncollected intervals
  std::vector<bool> ended_;        // whether the interval's end event was recorded
  std::vector<std::string> order_;
  std::vector<Acc> acc_;
};

}  // namespace alpacc
