// Launch-shape autotuner. Some kernels have launch shapes that all give
// bitwise the same results but run at different speeds depending on the
// grid and the device (the Cowan sweep's z-run length, the E update's
// workgroup size). For each (kernel, problem shape) key the tuner times its
// candidates with HIP events on their first uses in the real run, two
// samples each, and launches the fastest from then on. Decisions are kept
// per process and, when APRL_TUNE_CACHE names a file, read from it at the
// first use and appended to it when taken, so later runs start tuned.
#pragma once

#include <string>
#include <vector>

#include "common.hpp"

namespace aprl_gpu {

class TuneScope {
 public:
  // Picks the launch shape for `key` among `candidates` (the first is the
  // default, used when tuning is off); while the key is being tuned the
  // launch enqueued on `stream` within this scope is timed, unless `timed`
  // is false (a launch that may skip its work): it then takes the decision
  // if one is taken, else the default.
  TuneScope(const std::string& key, const std::vector<int>& candidates, hipStream_t stream, bool timed = true);
  ~TuneScope();
  TuneScope(const TuneScope&) = delete;
  TuneScope& operator=(const TuneScope&) = delete;
  int value() const { return value_; }

 private:
  int value_ = 0;
  void* entry_ = nullptr;
  int index_ = -1;
  hipEvent_t start_ = nullptr, stop_ = nullptr;
  hipStream_t stream_ = nullptr;
};

// "kernel NXxNYxNZ fp64|fp32": the key of a grid kernel
inline std::string shapeKey(const char* kernel, const aprl_grid_geometry& g, std::size_t word) {
  return std::string(kernel) + " " + std::to_string(g.nx) + "x" + std::to_string(g.ny) + "x" + std::to_string(g.nz) +
         (word == 8 ? " fp64" : " fp32");
}

// Tuning on or off for the process (on by default); off launches every
// key's first candidate.
void setAutotune(bool enabled);
// Whether `key` launches a settled shape (tuning off, or a decision taken
// or read from the cache), so that an untimed launch loses no sample.
bool autotuneSettled(const std::string& key);
// The decisions taken so far, one "key value" line each.
std::string autotuneReport();

}  // namespace aprl_gpu
