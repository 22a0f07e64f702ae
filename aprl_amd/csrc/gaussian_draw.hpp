// Seeded Gaussian draws of ParticleInitializer (particle_initializer.cpp:
// 24-50) and the light-front boost of a proper velocity, shared by the
// device beam generators. Every draw depends only on (seed, global particle
// index, component).
#pragma once

#include <hip/hip_runtime.h>

namespace aprl_gpu {
namespace draw {

constexpr double kPiInit = 3.141592653589793238462643383279502884;

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long v) {
  v += 0x9e3779b97f4a7c15ull;
  v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ull;
  v = (v ^ (v >> 27)) * 0x94d049bb133111ebull;
  return v ^ (v >> 31);
}

// uniformOpen / gaussian of particle_initializer.cpp:33-50.
__device__ __forceinline__ double uniformOpen(unsigned long long seed, unsigned long long particle,
                                              unsigned component) {
  const unsigned long long mixed = splitmix64(seed ^ splitmix64(particle + 0x100000001b3ull * component));
  const unsigned long long mantissa = (mixed >> 11) | 1ull;
  const double value = static_cast<double>(mantissa) * (1.0 / 9007199254740992.0);
  return fmin(value, 1.0 - 2.220446049250313e-16);
}

__device__ __forceinline__ double gaussianDraw(unsigned long long seed, unsigned long long particle, unsigned pair) {
  const double first = uniformOpen(seed, particle, 2 * pair);
  const double second = uniformOpen(seed, particle, 2 * pair + 1);
  return sqrt(-2.0 * log(first)) * cos(2.0 * kPiInit * second);
}

// BoostFrameTransform::gammaFromProperVelocity (boostframe.cpp:74-87).
__device__ __forceinline__ double gammaOf(const double u[3]) {
  const double scale = fmax(1.0, fmax(fabs(u[0]), fmax(fabs(u[1]), fabs(u[2]))));
  const double inv = 1.0 / scale;
  const double ux = u[0] * inv, uy = u[1] * inv, uz = u[2] * inv;
  return scale * sqrt(inv * inv + ux * ux + uy * uy + uz * uz);
}

// properVelocityLabToBox through the positive light-front component
// (boostframe.cpp:110-153); rapidity = gamma_b + gamma_b beta_b.
__device__ __forceinline__ void labToBox(double u[3], double rapidity) {
  const double gamma = gammaOf(u);
  const double mt2 = fma(u[0], u[0], fma(u[1], u[1], 1.0));
  const double positive = (u[2] >= 0.0 ? gamma + u[2] : mt2 / (gamma - u[2])) / rapidity;
  u[2] = 0.5 * (positive - mt2 / positive);
}

}  // namespace draw
}  // namespace aprl_gpu
