// Seeded Gaussian beam drawn on the device into a particle store: the
// generated-gaussian input of ParticleInitializer (particle_initializer.cpp:
// 24-100), one lane per macro-particle. Every draw depends only on (seed,
// global particle index, component), so the lanes are independent.
#include <cmath>

#include "gaussian_draw.hpp"
#include "handles.hpp"

namespace aprl_gpu {
namespace {

using namespace draw;

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
