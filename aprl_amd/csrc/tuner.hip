// Launch-shape autotuner (declared in tuner.hpp).
#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

#include "tuner.hpp"

namespace aprl_gpu {
namespace {

constexpr int kSamples = 2;  // timed launches per candidate

struct Pending {
  int index;
  hipEvent_t start, stop;
};

struct Entry {
  std::vector<int> candidates;
  std::vector<std::vector<float>> times;  // ms per timed launch, per candidate
  std::vector<Pending> pending;
  int chosen = -1;  // index into candidates
};

struct Registry {
  std::mutex lock;
  bool enabled = true;
  bool loaded = false;
  std::map<std::string, int> cached;  // decisions read from the cache file
  std::map<std::string, Entry> entries;
  std::vector<hipEvent_t> spare;

  hipEvent_t event() {
    if (!spare.empty()) {
      hipEvent_t e = spare.back();
      spare.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    hipCheck(hipEventCreate(&e), "hipEventCreate");
    return e;
  }

  static const char* cachePath() { return std::getenv("APRL_TUNE_CACHE"); }

  void load() {
    if (loaded) return;
    loaded = true;
    const char* path = cachePath();
    if (!path || !*path) return;
    std::ifstream in(path);
    std::string line;
    while (std::getline(in, line)) {
      const std::size_t cut = line.rfind(' ');
      if (cut == std::string::npos) continue;
      try {
        cached[line.substr(0, cut)] = std::stoi(line.substr(cut + 1));
      } catch (const std::exception&) {
        // a malformed line is ignored: the key is tuned again
      }
    }
  }

  void persist(const std::string& key, int value) {
    const char* path = cachePath();
    if (!path || !*path) return;
    std::ofstream out(path, std::ios::app);
    out << key << ' ' << value << '\n';
  }

  // Timings whose launches have finished; the decision once every
  // candidate has kSamples of them (the fastest by its slower sample).
  void harvest(const std::string& key, Entry& e) {
    for (std::size_t i = 0; i < e.pending.size();) {
      Pending& p = e.pending[i];
      if (hipEventQuery(p.stop) != hipSuccess) {
        ++i;
        continue;
      }
      float ms = 0.f;
      hipCheck(hipEventElapsedTime(&ms, p.start, p.stop), "hipEventElapsedTime");
      e.times[static_cast<std::size_t>(p.index)].push_back(ms);
      spare.push_back(p.start);
      spare.push_back(p.stop);
      e.pending.erase(e.pending.begin() + static_cast<std::ptrdiff_t>(i));
    }
    if (e.chosen >= 0) return;
    float best = 0.f;
    for (std::size_t c = 0; c < e.candidates.size(); ++c) {
      if (e.times[c].size() < static_cast<std::size_t>(kSamples)) return;
      const float t = *std::max_element(e.times[c].begin(), e.times[c].end());
      if (e.chosen < 0 || t < best) {
        best = t;
        e.chosen = static_cast<int>(c);
      }
    }
    persist(key, e.candidates[static_cast<std::size_t>(e.chosen)]);
  }
};

Registry& registry() {
  static Registry r;
  return r;
}

}  // namespace

TuneScope::TuneScope(const std::string& key, const std::vector<int>& candidates, hipStream_t stream, bool timed) {
  if (candidates.empty()) throw std::logic_error("Autotuner key without candidates");
  value_ = candidates.front();
  Registry& r = registry();
  std::lock_guard<std::mutex> guard(r.lock);
  if (!r.enabled || candidates.size() == 1) return;
  r.load();
  Entry& e = r.entries[key];
  if (e.candidates != candidates) {
    e = Entry{};
    e.candidates = candidates;
    e.times.assign(candidates.size(), {});
    const auto c = r.cached.find(key);
    if (c != r.cached.end()) {
      const auto it = std::find(candidates.begin(), candidates.end(), c->second);
      if (it != candidates.end()) e.chosen = static_cast<int>(it - candidates.begin());
    }
  }
  r.harvest(key, e);
  if (e.chosen >= 0) {
    value_ = e.candidates[static_cast<std::size_t>(e.chosen)];
    return;
  }
  if (!timed) return;  // the default, untimed
  // the candidate with the fewest timed or pending launches
  std::vector<std::size_t> uses(candidates.size());
  for (std::size_t c = 0; c < candidates.size(); ++c) uses[c] = e.times[c].size();
  for (const Pending& p : e.pending) ++uses[static_cast<std::size_t>(p.index)];
  index_ = static_cast<int>(std::min_element(uses.begin(), uses.end()) - uses.begin());
  if (uses[static_cast<std::size_t>(index_)] >= static_cast<std::size_t>(kSamples)) {
    index_ = -1;  // every sample in flight: launch the default untimed
    return;
  }
  value_ = candidates[static_cast<std::size_t>(index_)];
  entry_ = &e;
  stream_ = stream;
  start_ = r.event();
  stop_ = r.event();
  hipCheck(hipEventRecord(start_, stream), "hipEventRecord");
}

TuneScope::~TuneScope() {
  if (!entry_) return;
  Registry& r = registry();
  std::lock_guard<std::mutex> guard(r.lock);
  if (hipEventRecord(stop_, stream_) != hipSuccess) {
    r.spare.push_back(start_);
    r.spare.push_back(stop_);
    return;
  }
  static_cast<Entry*>(entry_)->pending.push_back(Pending{index_, start_, stop_});
}

void setAutotune(bool enabled) {
  Registry& r = registry();
  std::lock_guard<std::mutex> guard(r.lock);
  r.enabled = enabled;
}

bool autotuneSettled(const std::string& key) {
  Registry& r = registry();
  std::lock_guard<std::mutex> guard(r.lock);
  if (!r.enabled) return true;
  r.load();
  const auto it = r.entries.find(key);
  if (it == r.entries.end()) return false;  // not used yet: its first launches are timed
  r.harvest(key, it->second);
  return it->second.chosen >= 0;
}

std::string autotuneReport() {
  Registry& r = registry();
  std::lock_guard<std::mutex> guard(r.lock);
  std::ostringstream out;
  for (auto& kv : r.entries) {
    r.harvest(kv.first, kv.second);
    if (kv.second.chosen >= 0)
      out << kv.first << ' ' << kv.second.candidates[static_cast<std::size_t>(kv.second.chosen)] << '\n';
  }
  return out.str();
}

}  // namespace aprl_gpu

extern "C" {

aprl_status aprl_autotune(int enable) {
  return aprl_gpu::guarded([&] { aprl_gpu::setAutotune(enable != 0); });
}

aprl_status aprl_autotune_report(char* buffer, size_t capacity, size_t* length) {
  return aprl_gpu::guarded([&] {
    const std::string text = aprl_gpu::autotuneReport();
    if (length) *length = text.size();
    if (buffer && capacity) {
      const std::size_t n = std::min(capacity - 1, text.size());
      std::copy(text.begin(), text.begin() + static_cast<std::ptrdiff_t>(n), buffer);
      buffer[n] = '\0';
    }
  });
}

}  // extern "C"
