// Shared host/device utilities of the gfx950 library: status mapping,
// device buffers, per-kernel event timing.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2. Work items next to each other in a list (chunks of one tile box,
// blocks of neighbouring z planes) share data, so block b is given item
// (b mod 8)'s contiguous share of the list instead of item b: neighbouring
// items then run on one XCD and re-read its L2. A bijection on [0, n).
__device__ __forceinline__ int xcdContiguous(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

#include "aprl_gpu.h"

namespace aprl_gpu {

// SI constants of the reference (eb_field.cpp:15-19).
constexpr double kC = 299792458.0;
constexpr double kEps0 = 8.8541878128e-12;
constexpr double kMu0 = 1.25663706212e-6;
constexpr double kQe = 1.602176634e-19;
constexpr double kMe = 9.1093837139e-31;

struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void setLastError(const std::string& message);

inline void hipCheck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

// Workgroups of `kernel` (threads per block, no dynamic LDS) that are
// resident on the current device at once: compute units x the occupancy
// limit, queried once per kernel and device.
inline int residentBlocks(const void* kernel, int threads) {
  int dev = 0;
  hipCheck(hipGetDevice(&dev), "hipGetDevice");
  static std::mutex lock;
  static std::map<std::pair<const void*, int>, int> cache;
  std::lock_guard<std::mutex> guard(lock);
  const auto key = std::make_pair(kernel, dev);
  const auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int cus = 0, perCu = 0;
  hipCheck(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
  hipCheck(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, threads, 0),
           "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  const int r = cus * perCu > 0 ? cus * perCu : 1;
  cache[key] = r;
  return r;
}

// Run f and map the reference's exception types onto aprl_status.
template <class F>
aprl_status guarded(F&& f) {
  try {
    f();
    return APRL_OK;
  } catch (const DeviceError& e) {
    setLastError(e.what());
    return APRL_DEVICE_ERROR;
  } catch (const std::invalid_argument& e) {
    setLastError(e.what());
    return APRL_INVALID_ARGUMENT;
  } catch (const std::out_of_range& e) {
    setLastError(e.what());
    return APRL_OUT_OF_RANGE;
  } catch (const std::logic_error& e) {
    setLastError(e.what());
    return APRL_LOGIC_ERROR;
  } catch (const std::runtime_error& e) {
    setLastError(e.what());
    return APRL_RUNTIME_ERROR;
  } catch (const std::exception& e) {
    setLastError(e.what());
    return APRL_RUNTIME_ERROR;
  }
}

// Owning device allocation.
template <class T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(std::size_t n) { resize(n); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  ~DeviceArray() { release(); }
  void resize(std::size_t n) {
    release();
    if (n) hipCheck(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)), "hipMalloc");
    n_ = n;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  std::size_t size() const { return n_; }
  void swap(DeviceArray& o) noexcept { std::swap(p_, o.p_); std::swap(n_, o.n_); }

 private:
  T* p_ = nullptr;
  std::size_t n_ = 0;
};

// Small page-locked host buffer for stream-ordered readbacks.
class PinnedHost {
 public:
  PinnedHost() = default;
  PinnedHost(const PinnedHost&) = delete;
  PinnedHost& operator=(const PinnedHost&) = delete;
  ~PinnedHost() {
    if (p_) (void)hipHostFree(p_);
  }
  void* get(std::size_t bytes) {
    if (!p_) hipCheck(hipHostMalloc(&p_, bytes < 256 ? 256 : bytes, hipHostMallocDefault), "hipHostMalloc");
    return p_;
  }

 private:
  void* p_ = nullptr;
};

// Readback slots that kernels write straight into coherent page-locked host
// memory: a kernel stores its values, fences at system scope and then stores
// a sequence number in one of the flag words; the host spins on that word
// (a stream-ordered copy and a stream synchronization cost a copy kernel and
// a wake-up each). Layout in 32-bit words: [0, 8) flags, [8, 512) values;
// doubles from byte 2048.
class HostMailbox {
 public:
  static constexpr std::size_t kBytes = 4096;
  HostMailbox() = default;
  HostMailbox(const HostMailbox&) = delete;
  HostMailbox& operator=(const HostMailbox&) = delete;
  ~HostMailbox() {
    if (host_) (void)hipHostFree(host_);
  }
  unsigned* host() {
    if (!host_) {
      void* p = nullptr;
      hipCheck(hipHostMalloc(&p, kBytes, hipHostMallocMapped | hipHostMallocCoherent), "hipHostMalloc");
      std::memset(p, 0, kBytes);
      void* d = nullptr;
      hipCheck(hipHostGetDevicePointer(&d, p, 0), "hipHostGetDevicePointer");
      host_ = static_cast<unsigned*>(p);
      dev_ = static_cast<unsigned*>(d);
    }
    return host_;
  }
  unsigned* device() {
    host();
    return dev_;
  }
  double* hostDoubles() { return reinterpret_cast<double*>(reinterpret_cast<unsigned char*>(host()) + 2048); }
  // the next sequence number of flag word f
  unsigned next(int f) { return ++seq_[f]; }
  // Wait until flag word f holds seq; after a long wait the stream is
  // synchronized (reporting a failed kernel) and the flag checked again.
  void wait(int f, unsigned seq, hipStream_t st) {
    volatile unsigned* flag = host() + f;
    const auto start = std::chrono::steady_clock::now();
    for (unsigned spins = 0; *flag != seq; ++spins) {
      if ((spins & 1023u) == 1023u && std::chrono::steady_clock::now() - start > std::chrono::milliseconds(200)) {
        hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
        if (*flag != seq) throw DeviceError("A kernel readback did not arrive in host memory");
        break;
      }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
  }

 private:
  unsigned* host_ = nullptr;
  unsigned* dev_ = nullptr;
  unsigned seq_[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// A timing-free HIP event, created on first use.
class Event {
 public:
  Event() = default;
  Event(const Event&) = delete;
  Event& operator=(const Event&) = delete;
  ~Event() {
    if (e_) (void)hipEventDestroy(e_);
  }
  hipEvent_t get() {
    if (!e_) hipCheck(hipEventCreateWithFlags(&e_, hipEventDisableTiming), "hipEventCreate");
    return e_;
  }

 private:
  hipEvent_t e_ = nullptr;
};

// A non-blocking HIP stream at the default priority, created on first use.
class Stream {
 public:
  Stream() = default;
  Stream(const Stream&) = delete;
  Stream& operator=(const Stream&) = delete;
  ~Stream() {
    if (s_) (void)hipStreamDestroy(s_);
  }
  hipStream_t get() {
    if (!s_) hipCheck(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking), "hipStreamCreate");
    return s_;
  }

 private:
  hipStream_t s_ = nullptr;
};

// Per-kernel HIP-event timing on one stream. Events are recorded around each
// launch only while enabled; durations are collected lazily on query.
class Profiler {
 public:
  bool enabled = false;
  struct Pending { std::string name; hipEvent_t a, b; };
  struct Stat { uint64_t launches = 0; double ms = 0.0; };

  void begin(hipStream_t s, const char* name) {
    if (!enabled) return;
    Pending p;
    p.name = name;
    hipCheck(hipEventCreate(&p.a), "hipEventCreate");
    hipCheck(hipEventCreate(&p.b), "hipEventCreate");
    hipCheck(hipEventRecord(p.a, s), "hipEventRecord");
    open_.push_back(p);
  }
  void end(hipStream_t s) {
    if (!enabled) return;
    Pending& p = open_.back();
    hipCheck(hipEventRecord(p.b, s), "hipEventRecord");
    done_.push_back(p);
    open_.pop_back();
  }
  void collect() {
    for (auto& p : done_) {
      hipCheck(hipEventSynchronize(p.b), "hipEventSynchronize");
      float ms = 0.f;
      hipCheck(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
      Stat& s = stats_[p.name];
      ++s.launches;
      s.ms += ms;
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    done_.clear();
  }
  Stat query(const std::string& name) {
    collect();
    auto it = stats_.find(name);
    return it == stats_.end() ? Stat() : it->second;
  }
  void reset() {
    collect();
    stats_.clear();
  }
  ~Profiler() {
    for (auto& p : done_) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
  }

 private:
  std::vector<Pending> open_, done_;
  std::map<std::string, Stat> stats_;
};

struct ScopedKernel {
  Profiler& p;
  hipStream_t s;
  ScopedKernel(Profiler& prof, hipStream_t stream, const char* name) : p(prof), s(stream) { p.begin(s, name); }
  ~ScopedKernel() noexcept(false) { p.end(s); }
};

inline void launchCheck(const char* name) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw DeviceError(std::string("launch of ") + name + ": " + hipGetErrorString(e));
}

inline unsigned ceilDiv(std::size_t a, std::size_t b) { return static_cast<unsigned>((a + b - 1) / b); }

}  // namespace aprl_gpu
