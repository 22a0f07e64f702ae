// Seeded Gaussian beam drawn on the device into a particle store: the
// generated-gaussian input of ParticleInitializer (particle_initializer.cpp:
// 24-100), one lane per macro-particle. Every draw depends only on (seed,
// global particle index, component), so the lanes are independent.
#include <cmath>

#include "handles.hpp"

namespace aprl_gpu {
namespace {

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

struct BeamArgs {
  double center[3], sigmaX[3], meanU[3], sigmaU[3];
  double charge, mass, rapidity;
  unsigned long long seed, first;
  int boost;
};

__global__ void __launch_bounds__(256) kGaussianBeam(BeamArgs B, double* x, double* y, double* z, double* ux,
                                                     double* uy, double* uz, double* q, double* m, double* w,
                                                     uint64_t* id, int* state, double* beta, int* cell0,
                                                     int* cell1, int* cell2, float* f[9], float* fbeta, double hx,
                                                     double hy, double hz, long long n, unsigned* err) {
  const long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (k >= n) return;
  const unsigned long long global = B.first + static_cast<unsigned long long>(k);
  double pos[3], u[3];
  for (unsigned a = 0; a < 3; ++a) {
    pos[a] = B.center[a] + B.sigmaX[a] * gaussianDraw(B.seed, global, a);
    u[a] = B.meanU[a] + B.sigmaU[a] * gaussianDraw(B.seed ^ 0xd6e8feb86659fd93ull, global, a);
  }
  if (B.boost) labToBox(u, B.rapidity);
  if (x) {  // plain FP64 store
    x[k] = pos[0];
    y[k] = pos[1];
    z[k] = pos[2];
    ux[k] = u[0];
    uy[k] = u[1];
    uz[k] = u[2];
    q[k] = B.charge;
    m[k] = B.mass;
    w[k] = 1.0;
    beta[k] = 0.0;
  } else {  // FP32-split store: the host upload's splitCoordinate per axis
    const double h[3] = {hx, hy, hz};
    int* cells[3] = {cell0, cell1, cell2};
    for (int a = 0; a < 3; ++a) {
      const double r = pos[a] / h[a];
      if (!isfinite(r) || fabs(r) >= 2.0e9) {
        atomicOr(err, 1u);
        return;
      }
      double fl = floor(r);
      float off = static_cast<float>(r - fl);
      if (off >= 1.0f) {
        off = 0.0f;
        fl += 1.0;
      }
      cells[a][k] = static_cast<int>(fl);
      f[a][k] = off;
      f[3 + a][k] = static_cast<float>(u[a]);
    }
    f[6][k] = static_cast<float>(B.charge);
    f[7][k] = static_cast<float>(B.mass);
    f[8][k] = 1.0f;
    fbeta[k] = 0.0f;
  }
  id[k] = global + 1ull;
  state[k] = 0;
}

}  // namespace
}  // namespace aprl_gpu

using namespace aprl_gpu;

extern "C" aprl_status aprl_particles_generate_gaussian(aprl_particles* particles, const aprl_gaussian_beam* beam,
                                                        size_t count) {
  return guarded([&] {
    if (!particles || !beam) throw std::invalid_argument("Null particle store or beam");
    if (!(beam->total > 0)) throw std::invalid_argument("generated-gaussian requires macroparticles > 0");
    if (beam->first + count > beam->total) throw std::invalid_argument("Beam slice exceeds the beam's macroparticles");
    if (!(beam->electrons > 0.0) || !std::isfinite(beam->electrons))
      throw std::invalid_argument("Beam electrons must be positive and finite");
    for (int a = 0; a < 3; ++a)
      if (!std::isfinite(beam->center[a]) || !std::isfinite(beam->sigma_position[a]) ||
          !std::isfinite(beam->mean_u[a]) || !std::isfinite(beam->sigma_u[a]))
        throw std::invalid_argument("Beam parameters must be finite");
    if (!(beam->boost_gamma >= 1.0) || !std::isfinite(beam->boost_gamma))
      throw std::invalid_argument("Boost gamma must be finite and >= 1");
    BeamArgs B{};
    for (int a = 0; a < 3; ++a) {
      B.center[a] = beam->center[a];
      B.sigmaX[a] = beam->sigma_position[a];
      B.meanU[a] = beam->mean_u[a];
      B.sigmaU[a] = beam->sigma_u[a];
    }
    // particle_initializer.cpp:69-70, 94-95
    const double macro = beam->electrons / static_cast<double>(beam->total);
    B.charge = -kQe * macro;
    B.mass = kMe * macro;
    B.seed = beam->seed;
    B.first = beam->first;
    B.boost = beam->boost_gamma > 1.0 ? 1 : 0;
    B.rapidity = beam->boost_gamma + std::sqrt(beam->boost_gamma - 1.0) * std::sqrt(beam->boost_gamma + 1.0);
    hipStream_t st = particles->stream();
    hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
    DeviceArray<unsigned> err(1);
    hipCheck(hipMemsetAsync(err.get(), 0, sizeof(unsigned), st), "hipMemsetAsync");
    const long long n = static_cast<long long>(count);
    const unsigned blocks = ceilDiv(count ? count : 1, 256);
    if (particles->fields->fp32) {
      SplitStore& P = particles->sp;
      const FieldState32& F = particles->fields->s32;
      reserveSplitParticles(P, count);
      DeviceArray<float*> fptr(9);
      float* host[9];
      for (int a = 0; a < 9; ++a) host[a] = P.s[a].get();
      hipCheck(hipMemcpyAsync(fptr.get(), host, sizeof(host), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
      if (n)
        kGaussianBeam<<<blocks, 256, 0, st>>>(B, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                              nullptr, nullptr, P.id.get(), P.state.get(), nullptr, P.cell[0].get(),
                                              P.cell[1].get(), P.cell[2].get(), fptr.get(), P.beta.get(), F.g.dx,
                                              F.g.dy, F.g.dz, n, err.get());
      launchCheck("gaussian beam");
      hipCheck(hipMemsetAsync(P.removed.get(), 0, sizeof(unsigned), st), "hipMemsetAsync");
      unsigned e = 0;
      hipCheck(hipMemcpyAsync(&e, err.get(), sizeof(unsigned), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
      hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
      if (e) throw std::out_of_range("Particle position is outside the FP32-split lattice range");
      P.count = count;
      ++P.generation;
    } else {
      ParticleStore& P = particles->p;
      reserveParticles(P, count);
      if (n)
        kGaussianBeam<<<blocks, 256, 0, st>>>(B, P.s[0].get(), P.s[1].get(), P.s[2].get(), P.s[3].get(),
                                              P.s[4].get(), P.s[5].get(), P.s[6].get(), P.s[7].get(), P.s[8].get(),
                                              P.id.get(), P.state.get(), P.beta.get(), nullptr, nullptr, nullptr,
                                              nullptr, nullptr, 0.0, 0.0, 0.0, n, err.get());
      launchCheck("gaussian beam");
      hipCheck(hipMemsetAsync(P.removed.get(), 0, sizeof(unsigned), st), "hipMemsetAsync");
      hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
      P.count = count;
      ++P.generation;
    }
  });
}
