// Double-double (compensated) sums on the device: the accumulators stand in
// for the reference's long-double sums. Every block reduces its lanes in a
// fixed tree and the blocks are reduced in block order, so a sum is the
// same on every run.
#pragma once

#include <hip/hip_runtime.h>

namespace aprl_gpu {
namespace ddsum {

// double-double accumulator: value = hi + lo
struct DD {
  double hi = 0.0, lo = 0.0;
};
__device__ __forceinline__ DD ddAdd(DD a, double b) {
  const double s = a.hi + b, bb = s - a.hi, e = (a.hi - (s - bb)) + (b - bb) + a.lo;
  const double h = s + e;
  return DD{h, e - (h - s)};
}
__device__ __forceinline__ DD ddAdd(DD a, DD b) {
  const double s = a.hi + b.hi, bb = s - a.hi, e = (a.hi - (s - bb)) + (b.hi - bb) + a.lo + b.lo;
  const double h = s + e;
  return DD{h, e - (h - s)};
}
// exact product x * y added
__device__ __forceinline__ DD ddAddProduct(DD a, double x, double y) {
  const double p = x * y;
  return ddAdd(ddAdd(a, p), fma(x, y, -p));
}

// Block reduction of K accumulators into out[K * blockIdx.x + c] (fixed tree).
template <int K, int kThreads = 256>
__device__ void blockDD(DD v[K], DD* out) {
  __shared__ double sh[K][kThreads], sl[K][kThreads];
  for (int c = 0; c < K; ++c) {
    sh[c][threadIdx.x] = v[c].hi;
    sl[c][threadIdx.x] = v[c].lo;
  }
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w)
      for (int c = 0; c < K; ++c) {
        const DD r = ddAdd(DD{sh[c][threadIdx.x], sl[c][threadIdx.x]}, DD{sh[c][threadIdx.x + w], sl[c][threadIdx.x + w]});
        sh[c][threadIdx.x] = r.hi;
        sl[c][threadIdx.x] = r.lo;
      }
    __syncthreads();
  }
  if (threadIdx.x < static_cast<unsigned>(K)) out[K * blockIdx.x + threadIdx.x] = DD{sh[threadIdx.x][0], sl[threadIdx.x][0]};
}

namespace {
// out[c] = sum over blocks of in[K * b + c], in block order (one thread each).
__global__ void kFinishDD(const DD* in, int blocks, int K, DD* out) {
  const int c = threadIdx.x;
  if (c >= K) return;
  DD s;
  for (int b = 0; b < blocks; ++b) s = ddAdd(s, in[K * b + c]);
  out[c] = s;
}
}  // namespace

// Order-preserving 64-bit image of a double (for exact atomic min / max).
__device__ __forceinline__ unsigned long long orderedBits(double v) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
inline double fromOrderedBits(unsigned long long o) {
  const unsigned long long b = (o >> 63) ? (o & 0x7fffffffffffffffull) : ~o;
  double v;
  __builtin_memcpy(&v, &b, sizeof v);
  return v;
}
// Slot values before the first atomic: above every ordered image (min) and
// below every one (max).
constexpr unsigned long long kOrderedMinStart = 0xffffffffffffffffull;
constexpr unsigned long long kOrderedMaxStart = 0ull;

}  // namespace ddsum
}  // namespace aprl_gpu
