// FP64 diagnostics of the field grid and the coupled step on the device:
// interior field energy, outward Poynting power through the faces of a
// cell box (the energy-ledger channels), the charge-continuity residual of
// the last coupled step and the Gauss-law residual. Every sum is a
// compensated, fixed-order reduction, so results do not depend on timing.
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "diagnostics.hpp"
#include "charge_basis.hpp"
#include "particles.hpp"

namespace aprl_gpu {

namespace {

constexpr int kRedBlocks = 512, kRedThreads = 256;

// Neumaier two-term accumulator.
struct Acc {
  double s = 0.0, c = 0.0;
  __device__ void add(double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  __device__ void add(const Acc& o) {
    add(o.s);
    add(o.c);
  }
};

template <class R>
__device__ __forceinline__ double at(const R* p, const int* n, int i, int j, int k) {
  return static_cast<double>(p[(static_cast<size_t>(k) * n[1] + j) * n[0] + i]);
}

// EBFieldGrid::fieldEnergyDensityCell, eb_field.cpp:715-738.
template <class R>
__device__ double energyDensityCell(const FieldPtrsT<R>& F, int i, int j, int k) {
  const double e0 = 0.25 * (at(F.e[0], F.n[0], i, j, k) + at(F.e[0], F.n[0], i, j + 1, k) +
                            at(F.e[0], F.n[0], i, j, k + 1) + at(F.e[0], F.n[0], i, j + 1, k + 1));
  const double e1 = 0.25 * (at(F.e[1], F.n[1], i, j, k) + at(F.e[1], F.n[1], i + 1, j, k) +
                            at(F.e[1], F.n[1], i, j, k + 1) + at(F.e[1], F.n[1], i + 1, j, k + 1));
  const double e2 = 0.25 * (at(F.e[2], F.n[2], i, j, k) + at(F.e[2], F.n[2], i + 1, j, k) +
                            at(F.e[2], F.n[2], i, j + 1, k) + at(F.e[2], F.n[2], i + 1, j + 1, k));
  const double b0 = 0.5 * (at(F.b[0], F.n[3], i, j, k) + at(F.b[0], F.n[3], i + 1, j, k));
  const double b1 = 0.5 * (at(F.b[1], F.n[4], i, j, k) + at(F.b[1], F.n[4], i, j + 1, k));
  const double b2 = 0.5 * (at(F.b[2], F.n[5], i, j, k) + at(F.b[2], F.n[5], i, j, k + 1));
  return 0.5 * (kEps0 * (e0 * e0 + e1 * e1 + e2 * e2) + (b0 * b0 + b1 * b1 + b2 * b2) / kMu0);
}

// EBFieldGrid::poyntingComponentCell, eb_field.cpp:740-782.
template <class R>
__device__ double poyntingCell(const FieldPtrsT<R>& F, int i, int j, int k, int axis) {
  auto cEx = [&]() {
    return 0.25 * (at(F.e[0], F.n[0], i, j, k) + at(F.e[0], F.n[0], i, j + 1, k) + at(F.e[0], F.n[0], i, j, k + 1) +
                   at(F.e[0], F.n[0], i, j + 1, k + 1));
  };
  auto cEy = [&]() {
    return 0.25 * (at(F.e[1], F.n[1], i, j, k) + at(F.e[1], F.n[1], i + 1, j, k) + at(F.e[1], F.n[1], i, j, k + 1) +
                   at(F.e[1], F.n[1], i + 1, j, k + 1));
  };
  auto cEz = [&]() {
    return 0.25 * (at(F.e[2], F.n[2], i, j, k) + at(F.e[2], F.n[2], i + 1, j, k) + at(F.e[2], F.n[2], i, j + 1, k) +
                   at(F.e[2], F.n[2], i + 1, j + 1, k));
  };
  auto cBx = [&]() { return 0.5 * (at(F.b[0], F.n[3], i, j, k) + at(F.b[0], F.n[3], i + 1, j, k)); };
  auto cBy = [&]() { return 0.5 * (at(F.b[1], F.n[4], i, j, k) + at(F.b[1], F.n[4], i, j + 1, k)); };
  auto cBz = [&]() { return 0.5 * (at(F.b[2], F.n[5], i, j, k) + at(F.b[2], F.n[5], i, j, k + 1)); };
  double eA, eB, bA, bB;
  if (axis == 0) { eA = cEy(); eB = cEz(); bA = cBy(); bB = cBz(); }
  else if (axis == 1) { eA = cEz(); eB = cEx(); bA = cBz(); bB = cBx(); }
  else { eA = cEx(); eB = cEy(); bA = cBx(); bB = cBy(); }
  double s = eA * bB - eB * bA;
  s /= kMu0;
  return s;
}

// Block reduction of one accumulator in a fixed tree order; thread 0
// returns the block total.
__device__ Acc blockReduce(Acc a, double* ss, double* sc) {
  ss[threadIdx.x] = a.s;
  sc[threadIdx.x] = a.c;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) {
      Acc x{ss[threadIdx.x], sc[threadIdx.x]};
      x.add(Acc{ss[threadIdx.x + w], sc[threadIdx.x + w]});
      ss[threadIdx.x] = x.s;
      sc[threadIdx.x] = x.c;
    }
    __syncthreads();
  }
  return Acc{ss[0], sc[0]};
}

// Sum of energyDensityCell * volume over the cells of [lo, hi).
template <class R>
__global__ void __launch_bounds__(kRedThreads) kEnergyBox(FieldPtrsT<R> F, int3 lo, int3 hi, double volume,
                                                          double* partial) {
  __shared__ double ss[kRedThreads], sc[kRedThreads];
  const long long wx = hi.x - lo.x, wy = hi.y - lo.y, wz = hi.z - lo.z, total = wx * wy * wz;
  Acc a;
  for (long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; t < total;
       t += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int i = lo.x + static_cast<int>(t % wx), j = lo.y + static_cast<int>((t / wx) % wy),
              k = lo.z + static_cast<int>(t / (wx * wy));
    a.add(energyDensityCell(F, i, j, k) * volume);
  }
  const Acc b = blockReduce(a, ss, sc);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = b.s;
    partial[2 * blockIdx.x + 1] = b.c;
  }
}

// Outward power through the six faces of the cell box [lo, hi):
// face 2a (lower, sign -) and 2a+1 (upper, sign +) sum the axis-a Poynting
// component over the face's cells times the face area. blockIdx.y = face.
template <class R>
__global__ void __launch_bounds__(kRedThreads) kFacePower(FieldPtrsT<R> F, int3 lo, int3 hi, double3 area,
                                                          double* partial) {
  __shared__ double ss[kRedThreads], sc[kRedThreads];
  const int face = blockIdx.y, axis = face / 2;
  const bool upper = face & 1;
  const int l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
  const int ua = (axis + 1) % 3, va = (axis + 2) % 3;
  // loop order of the reference: the in-face axes run x fastest
  const int fa = ua < va ? ua : va, sa = ua < va ? va : ua;
  const long long wf = h[fa] - l[fa], ws = h[sa] - l[sa], total = wf * ws;
  const double ar = axis == 0 ? area.x : (axis == 1 ? area.y : area.z);
  Acc a;
  for (long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; t < total;
       t += static_cast<long long>(gridDim.x) * blockDim.x) {
    int c[3];
    c[axis] = upper ? h[axis] - 1 : l[axis];
    c[fa] = l[fa] + static_cast<int>(t % wf);
    c[sa] = l[sa] + static_cast<int>(t / wf);
    const double s = poyntingCell(F, c[0], c[1], c[2], axis) * ar;
    a.add(upper ? s : -s);
  }
  const Acc b = blockReduce(a, ss, sc);
  if (threadIdx.x == 0) {
    partial[2 * (face * gridDim.x + blockIdx.x)] = b.s;
    partial[2 * (face * gridDim.x + blockIdx.x) + 1] = b.c;
  }
}

// Final compensated sum of `blocks` partial accumulators per output.
__global__ void __launch_bounds__(kRedThreads) kFinish(const double* partial, int blocks, double* out) {
  __shared__ double ss[kRedThreads], sc[kRedThreads];
  const double* p = partial + 2 * static_cast<size_t>(blockIdx.x) * blocks;
  Acc a;
  for (int b = threadIdx.x; b < blocks; b += blockDim.x) a.add(Acc{p[2 * b], p[2 * b + 1]});
  const Acc r = blockReduce(a, ss, sc);
  if (threadIdx.x == 0) out[blockIdx.x] = r.s + r.c;
}

// ---------------------------------------------------------------------------
// Charge density, continuity and Gauss residuals.

struct RhoArgs {
  const double* x[3];  // absolute positions (plain form), or
  const int* cell[3];  // lattice cell and in-cell offset (FP32-split form)
  const float* off[3];
  int oc[3];           // lattice cell of the grid origin (split form)
  const double* q;
  const float* qf;     // charges of the split form
  int n;
  double o[3], h[3];
  int nx, ny, nz;
  int shape;  // 0 compensated quintic, 2 binomial3 with x-binomial3
  int shift;
  int absolute;  // 1: accumulate |tap| (the bound pass)
  double invVolume;
  unsigned long long* acc;  // (nx+1)(ny+1)(nz+1) vertices
  ChargeWeights cw;
};

// Shaped charge of every particle on the vertices, in 64-bit fixed point:
// binomial3-x: x nodeWeights (4) x linear y (2) x binomial3 z basis (4),
// the charge whose change the binomial3-x current conserves
// (verification/transverse_binomial_checks.cpp:33-40); quintic: linear x, y
// (2 x 2) x quintic z basis (12). Vertices outside the grid are dropped.
__global__ void kChargeDensity(RhoArgs R) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= R.n) return;
  double factor = 1.0;
  if (R.cw.mode == 1) {
    const int st = R.cw.state[p];
    if (st < 0) return;
    if (st == kStateCarrier) factor = R.cw.wd ? R.cw.wd[p] : static_cast<double>(R.cw.wf[p]);
  } else if (R.cw.mode == 2 && R.cw.evt[p]) {
    factor = R.cw.wd ? R.cw.wd[p] : static_cast<double>(R.cw.wf[p]);
  }
  double l[3];
  for (int a = 0; a < 3; ++a)
    l[a] = R.x[0] ? (R.x[a][p] - R.o[a]) / R.h[a]
                  : static_cast<double>(R.cell[a][p] - R.oc[a]) + static_cast<double>(R.off[a][p]);
  const double lx = l[0], ly = l[1], lz = l[2];
  const int i = static_cast<int>(floor(lx)), j = static_cast<int>(floor(ly));
  const double fx = lx - i, fy = ly - j;
  double wz[12];
  const int k0 = chargeBasis(R.shape, lz, wz);
  const int nz = R.shape == 2 ? 4 : 12;
  double wx[4];
  int nxw, i0;
  if (R.shape == 2) {
    wx[0] = 0.25 * (1. - fx); wx[1] = 0.5 - 0.25 * fx; wx[2] = 0.25 + 0.25 * fx; wx[3] = 0.25 * fx;
    nxw = 4;
    i0 = i - 1;
  } else {
    wx[0] = 1. - fx; wx[1] = fx;
    nxw = 2;
    i0 = i;
  }
  const double wy[2] = {1. - fy, fy};
  const double qv = (R.q ? R.q[p] : static_cast<double>(R.qf[p])) * factor * R.invVolume;
  const int vx = R.nx + 1, vy = R.ny + 1, vz = R.nz + 1;
  for (int c = 0; c < nz; ++c) {
    const int kk = k0 + c;
    if (kk < 0 || kk >= vz) continue;
    for (int b = 0; b < 2; ++b) {
      const int jj = j + b;
      if (jj < 0 || jj >= vy) continue;
      for (int a = 0; a < nxw; ++a) {
        const int ii = i0 + a;
        if (ii < 0 || ii >= vx) continue;
        const double w = qv * wx[a] * wy[b] * wz[c];
        atomicAdd(R.acc + (static_cast<size_t>(kk) * vy + jj) * vx + ii,
                  static_cast<unsigned long long>(llrint(ldexp(R.absolute ? fabs(w) : w, R.shift))));
      }
    }
  }
}

__global__ void kFixedToDouble(unsigned long long* acc, double* out, size_t n, int shift) {
  for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<size_t>(gridDim.x) * blockDim.x) {
    out[t] = ldexp(static_cast<double>(static_cast<long long>(acc[t])), -shift);
    acc[t] = 0ull;
  }
}

// Largest entry of a non-negative fixed-point array into *out (zeroed
// before); clears what it reads.
__global__ void __launch_bounds__(kRedThreads) kMaxClearU64(unsigned long long* acc, size_t n,
                                                             unsigned long long* out) {
  __shared__ unsigned long long sm[kRedThreads];
  unsigned long long m = 0ull;
  for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<size_t>(gridDim.x) * blockDim.x) {
    m = max(m, acc[t]);
    acc[t] = 0ull;
  }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) sm[threadIdx.x] = max(sm[threadIdx.x], sm[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, sm[0]);
}

template <class T>
__global__ void kSumAbs(const T* q, int n, double* partial) {
  __shared__ double ss[kRedThreads], sc[kRedThreads];
  Acc a;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) a.add(fabs(static_cast<double>(q[p])));
  const Acc b = blockReduce(a, ss, sc);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = b.s;
    partial[2 * blockIdx.x + 1] = b.c;
  }
}

__device__ __forceinline__ void atomicMaxAbs(unsigned long long* slot, double v) {
  // non-negative doubles order like their bit patterns
  atomicMax(slot, static_cast<unsigned long long>(__double_as_longlong(fabs(v))));
}

// ChargeConservingCurrentDepositor::maxContinuityResidual,
// eb_deposition.cpp:195-226: max over vertices of
// |(rho1 - rho0)/dt + div J|, and of |(rho1 - rho0)/dt| as its scale.
template <class R>
__global__ void kContinuity(FieldPtrsT<R> F, const double* rho0, const double* rho1, double dt, double3 h,
                            unsigned long long* out) {
  const int nx = F.n[0][0], ny = F.n[1][1], nz = F.n[2][2];
  const size_t total = static_cast<size_t>(nx + 1) * (ny + 1) * (nz + 1);
  for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < total;
       t += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int i = static_cast<int>(t % (nx + 1)), j = static_cast<int>((t / (nx + 1)) % (ny + 1)),
              k = static_cast<int>(t / (static_cast<size_t>(nx + 1) * (ny + 1)));
    const double rightX = i < nx ? at(F.j[0], F.n[6], i, j, k) : 0.0;
    const double leftX = i > 0 ? at(F.j[0], F.n[6], i - 1, j, k) : 0.0;
    const double rightY = j < ny ? at(F.j[1], F.n[7], i, j, k) : 0.0;
    const double leftY = j > 0 ? at(F.j[1], F.n[7], i, j - 1, k) : 0.0;
    const double rightZ = k < nz ? at(F.j[2], F.n[8], i, j, k) : 0.0;
    const double leftZ = k > 0 ? at(F.j[2], F.n[8], i, j, k - 1) : 0.0;
    const double rate = (rho1[t] - rho0[t]) / dt;
    const double residual = rate + (rightX - leftX) / h.x + (rightY - leftY) / h.y + (rightZ - leftZ) / h.z;
    atomicMaxAbs(out, residual);
    atomicMaxAbs(out + 1, rate);
    atomicMaxAbs(out + 2, fmax(fabs(rho0[t]), fabs(rho1[t])) / dt);
  }
}

// Gauss residual of eb_initial_field.cpp:774-799 over the vertices strictly
// inside [lo, hi]: error = div E - rho / eps0; per block the compensated
// sums of error^2 and source^2, and the max |error| and |source|.
template <class R>
__global__ void __launch_bounds__(kRedThreads) kGauss(FieldPtrsT<R> F, const double* rho, int3 lo, int3 hi, double3 h,
                                                      double* partial, unsigned long long* maxima) {
  __shared__ double ss[kRedThreads], sc[kRedThreads];
  const int nx = F.n[0][0], ny = F.n[1][1];
  const long long wx = hi.x - lo.x - 1, wy = hi.y - lo.y - 1, wz = hi.z - lo.z - 1;
  const long long total = wx > 0 && wy > 0 && wz > 0 ? wx * wy * wz : 0;
  Acc e2, s2;
  for (long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; t < total;
       t += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int i = lo.x + 1 + static_cast<int>(t % wx), j = lo.y + 1 + static_cast<int>((t / wx) % wy),
              k = lo.z + 1 + static_cast<int>(t / (wx * wy));
    const double divergence = (at(F.e[0], F.n[0], i, j, k) - at(F.e[0], F.n[0], i - 1, j, k)) / h.x +
                              (at(F.e[1], F.n[1], i, j, k) - at(F.e[1], F.n[1], i, j - 1, k)) / h.y +
                              (at(F.e[2], F.n[2], i, j, k) - at(F.e[2], F.n[2], i, j, k - 1)) / h.z;
    const double source = rho[(static_cast<size_t>(k) * (ny + 1) + j) * (nx + 1) + i] / kEps0;
    const double error = divergence - source;
    e2.add(error * error);
    s2.add(source * source);
    atomicMaxAbs(maxima, error);
    atomicMaxAbs(maxima + 1, source);
  }
  const Acc a = blockReduce(e2, ss, sc);
  __syncthreads();
  const Acc b = blockReduce(s2, ss, sc);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = a.s;
    partial[2 * blockIdx.x + 1] = a.c;
    partial[2 * (kRedBlocks + blockIdx.x)] = b.s;
    partial[2 * (kRedBlocks + blockIdx.x) + 1] = b.c;
  }
}

__global__ void kZeroU64(unsigned long long* p, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0ull;
}

template <class T>
void ensure(DeviceArray<T>& a, std::size_t n) {
  if (a.size() < n) a.resize(n);
}

int3 toInt3(const std::size_t v[3]) {
  return make_int3(static_cast<int>(v[0]), static_cast<int>(v[1]), static_cast<int>(v[2]));
}

template <class R>
void checkBox(const FieldStateT<R>& s, const std::size_t lo[3], const std::size_t hi[3], bool cells) {
  const std::size_t n[3] = {s.g.nx, s.g.ny, s.g.nz};
  for (int a = 0; a < 3; ++a)
    if (lo[a] > hi[a] || hi[a] > n[a] + (cells ? 0 : 1))
      throw std::out_of_range("Diagnostic box is outside the E/B grid");
}

}  // namespace

template <class R>
double interiorEnergy(FieldStateT<R>& s, const std::size_t lo[3], const std::size_t hi[3]) {
  checkBox(s, lo, hi, true);
  DiagnosticScratch& D = s.diag;
  ensure(D.partial, 2 * 6 * kRedBlocks);
  ensure(D.out, 8);
  const double volume = s.g.dx * s.g.dy * s.g.dz;
  kEnergyBox<<<kRedBlocks, kRedThreads, 0, s.stream>>>(s.ptrs(), toInt3(lo), toInt3(hi), volume, D.partial.get());
  kFinish<<<1, kRedThreads, 0, s.stream>>>(D.partial.get(), kRedBlocks, D.out.get());
  launchCheck("interior energy");
  double e = 0.0;
  hipCheck(hipMemcpyAsync(&e, D.out.get(), sizeof(double), hipMemcpyDeviceToHost, s.stream), "hipMemcpyAsync");
  hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
  return e;
}

template <class R>
void facePower(FieldStateT<R>& s, const std::size_t lo[3], const std::size_t hi[3], double power[6]) {
  checkBox(s, lo, hi, true);
  for (int a = 0; a < 3; ++a)
    if (lo[a] >= hi[a]) throw std::out_of_range("Diagnostic face box is empty");
  DiagnosticScratch& D = s.diag;
  ensure(D.partial, 2 * 6 * kRedBlocks);
  ensure(D.out, 8);
  const double3 area = make_double3(s.g.dy * s.g.dz, s.g.dx * s.g.dz, s.g.dx * s.g.dy);
  constexpr int kFaceBlocks = 64;
  kFacePower<<<dim3(kFaceBlocks, 6), kRedThreads, 0, s.stream>>>(s.ptrs(), toInt3(lo), toInt3(hi), area,
                                                                   D.partial.get());
  kFinish<<<6, kRedThreads, 0, s.stream>>>(D.partial.get(), kFaceBlocks, D.out.get());
  launchCheck("face power");
  hipCheck(hipMemcpyAsync(power, D.out.get(), 6 * sizeof(double), hipMemcpyDeviceToHost, s.stream),
           "hipMemcpyAsync");
  hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
}

// Shaped charge density of a particle set on the vertices into `rho`; R
// carries the positions and charges, q the charges for the bound.
template <class Real, class Q>
void depositCharge(FieldStateT<Real>& s, RhoArgs R, const Q* q, const RhoWorkspace& ws, double* rho) {
  const std::size_t nv = (s.g.nx + 1) * (s.g.ny + 1) * (s.g.nz + 1);
  // global fixed-point exponent from the total |charge| density bound (each
  // shape's weights sum to at most 4 in absolute value)
  const int n = R.n;
  double total = 0.0;
  if (n > 0) {
    kSumAbs<<<kRedBlocks, kRedThreads, 0, s.stream>>>(q, n, ws.partial);
    kFinish<<<1, kRedThreads, 0, s.stream>>>(ws.partial, kRedBlocks, ws.out);
    launchCheck("charge bound");
    hipCheck(hipMemcpyAsync(&total, ws.out, sizeof(double), hipMemcpyDeviceToHost, s.stream), "hipMemcpyAsync");
    hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
  }
  const double invVolume = 1.0 / (s.g.dx * s.g.dy * s.g.dz);
  R.h[0] = s.g.dx; R.h[1] = s.g.dy; R.h[2] = s.g.dz;
  R.nx = static_cast<int>(s.g.nx);
  R.ny = static_cast<int>(s.g.ny);
  R.nz = static_cast<int>(s.g.nz);
  R.invVolume = invVolume;
  R.acc = ws.acc;
  int shift = 0;
  if (n > 0 && total > 0.0) {
    // Bound pass: sum |tap| per vertex on the coarse exponent of the total
    // |charge| (no vertex exceeds it), and take the largest vertex. The
    // signed pass then uses the exponent of twice that bound, so its
    // quantum is 2^-61 of the largest |rho| rather than of the whole
    // bunch's charge (which for a diffuse bunch of many particles is
    // orders of magnitude larger than any one vertex).
    int e2 = 0;
    (void)std::frexp(4.0 * total * invVolume, &e2);
    const int coarse = 62 - e2;
    R.shift = coarse;
    R.absolute = 1;
    kZeroU64<<<1, 64, 0, s.stream>>>(ws.maxima, 4);
    kChargeDensity<<<ceilDiv(n, 256), 256, 0, s.stream>>>(R);
    kMaxClearU64<<<kRedBlocks, kRedThreads, 0, s.stream>>>(ws.acc, nv, ws.maxima);
    launchCheck("charge bound");
    unsigned long long peak = 0ull;
    hipCheck(hipMemcpyAsync(&peak, ws.maxima, sizeof(peak), hipMemcpyDeviceToHost, s.stream), "hipMemcpyAsync");
    hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
    // Each tap rounds by at most half a coarse quantum: negligible beside
    // the peak (the taps of a particle sum to at least its |charge|, so the
    // peak holds at least 2^62 / (8 vertices) quanta), and covered by the
    // factor 2 of headroom below.
    const double bound = std::ldexp(static_cast<double>(peak) + 1.0, -coarse);
    (void)std::frexp(2.0 * bound, &e2);
    shift = 62 - e2;
  }
  R.shift = shift;
  R.absolute = 0;
  if (n > 0) kChargeDensity<<<ceilDiv(n, 256), 256, 0, s.stream>>>(R);
  kFixedToDouble<<<2048, 256, 0, s.stream>>>(ws.acc, rho, nv, shift);
  launchCheck("charge density");
}

template <class R>
RhoWorkspace diagnosticWorkspace(FieldStateT<R>& s) {
  DiagnosticScratch& D = s.diag;
  const std::size_t nv = (s.g.nx + 1) * (s.g.ny + 1) * (s.g.nz + 1);
  if (D.acc.size() != nv) {
    D.acc.resize(nv);
    hipCheck(hipMemsetAsync(D.acc.get(), 0, nv * sizeof(unsigned long long), s.stream), "hipMemsetAsync");
  }
  ensure(D.partial, 2 * 6 * kRedBlocks);
  ensure(D.out, 8);
  ensure(D.maxima, 4);
  RhoWorkspace ws;
  ws.acc = D.acc.get();
  ws.partial = D.partial.get();
  ws.out = D.out.get();
  ws.maxima = D.maxima.get();
  return ws;
}

void chargeDensity(FieldState& s, int shape, const double* const x[3], const double* q, int n, const double origin[3],
                   const ChargeWeights& cw, const RhoWorkspace& ws, double* rho) {
  RhoArgs R{};
  R.cw = cw;
  for (int a = 0; a < 3; ++a) {
    R.x[a] = x[a];
    R.o[a] = origin[a];
  }
  R.q = q;
  R.n = n;
  R.shape = shape;
  depositCharge(s, R, q, ws, rho);
}

void chargeDensitySplit(FieldState32& s, int shape, const int* const cell[3], const float* const off[3],
                        const float* q, int n, const int oc[3], const ChargeWeights& cw, const RhoWorkspace& ws,
                        double* rho) {
  RhoArgs R{};
  R.cw = cw;
  for (int a = 0; a < 3; ++a) {
    R.cell[a] = cell[a];
    R.off[a] = off[a];
    R.oc[a] = oc[a];
  }
  R.qf = q;
  R.n = n;
  R.shape = shape;
  depositCharge(s, R, q, ws, rho);
}

template <class R>
void continuityResidual(FieldStateT<R>& s, const double* rho0, const double* rho1, const RhoWorkspace& ws,
                        double out[3]) {
  kZeroU64<<<1, 64, 0, s.stream>>>(ws.maxima, 4);
  kContinuity<<<2048, 256, 0, s.stream>>>(s.ptrs(), rho0, rho1, s.g.dt, make_double3(s.g.dx, s.g.dy, s.g.dz),
                                          ws.maxima);
  launchCheck("continuity");
  unsigned long long m[3];
  hipCheck(hipMemcpyAsync(m, ws.maxima, sizeof(m), hipMemcpyDeviceToHost, s.stream), "hipMemcpyAsync");
  hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
  std::memcpy(out, m, sizeof(m));
}

// ---------------------------------------------------------------------------
// Allocation-free continuity workspace: one file-scope device array, laid
// out as the fixed-point accumulator, the two charge densities (one word
// per vertex each, up to kNoAllocVertices) and the reduction partials.
// Every GPU the library runs on holds its own copy for the life of the
// process; its device address comes from hipGetSymbolAddress, taken once
// per device with that device current (the array's name is not a device
// address in host code). Calls on one GPU share it: overlapping calls race.

namespace {
constexpr std::size_t kNoAllocVertices = 2400000;  // 64 x 64 x 512 cells have 2,167,425 vertices
constexpr std::size_t kNoAllocPartial = 2 * kRedBlocks, kNoAllocOut = 8, kNoAllocMaxima = 4;
constexpr std::size_t kNoAllocWords = 3 * kNoAllocVertices + kNoAllocPartial + kNoAllocOut + kNoAllocMaxima;
}  // namespace

static __device__ __align__(256) unsigned long long scratch[kNoAllocWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
unsigned long long* gScratch[kMaxDevices] = {};

__global__ void kFillU64(unsigned long long* p, size_t n) {
  for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[t] = 0ull;
}

unsigned long long* scratchOnCurrentDevice() {
  int device = 0;
  hipCheck(hipGetDevice(&device), "hipGetDevice");
  if (device < 0 || device >= kMaxDevices) throw std::runtime_error("Allocation-free workspace: device index out of range");
  std::lock_guard<std::mutex> lock(gScratchMutex);
  if (!gScratch[device]) {
    void* address = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&address, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !address)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(rc));
    gScratch[device] = static_cast<unsigned long long*>(address);
  }
  return gScratch[device];
}
}  // namespace

std::size_t noAllocVertexLimit() { return kNoAllocVertices; }

template <class R>
RhoWorkspace noAllocWorkspace(FieldStateT<R>& s, std::size_t vertices, double** rho0, double** rho1) {
  if (vertices > kNoAllocVertices)
    throw std::invalid_argument("Allocation-free continuity: the grid has " + std::to_string(vertices) +
                            " vertices, the workspace holds at most " + std::to_string(kNoAllocVertices));
  unsigned long long* base = scratchOnCurrentDevice();
  RhoWorkspace ws;
  ws.acc = base;
  *rho0 = reinterpret_cast<double*>(base + kNoAllocVertices);
  *rho1 = reinterpret_cast<double*>(base + 2 * kNoAllocVertices);
  ws.partial = reinterpret_cast<double*>(base + 3 * kNoAllocVertices);
  ws.out = ws.partial + kNoAllocPartial;
  ws.maxima = base + 3 * kNoAllocVertices + kNoAllocPartial + kNoAllocOut;
  // the accumulator keeps the last call's contents; clear the part in use
  if (vertices) kFillU64<<<static_cast<unsigned>(std::min<std::size_t>(ceilDiv(vertices, 256), 2048)), 256, 0,
                           s.stream>>>(ws.acc, vertices);
  launchCheck("allocation-free workspace clear");
  return ws;
}

template <class R>
void gaussResidual(FieldStateT<R>& s, const DeviceArray<double>& rho, const std::size_t lo[3], const std::size_t hi[3],
                   double out[4]) {
  // strictly interior vertices i in (lo, hi) read Ex(i), Ey(j), Ez(k): hi <= n
  checkBox(s, lo, hi, true);
  DiagnosticScratch& D = s.diag;
  ensure(D.partial, 2 * 6 * kRedBlocks);
  ensure(D.out, 8);
  ensure(D.maxima, 4);
  kZeroU64<<<1, 64, 0, s.stream>>>(D.maxima.get(), 4);
  kGauss<<<kRedBlocks, kRedThreads, 0, s.stream>>>(s.ptrs(), rho.get(), toInt3(lo), toInt3(hi),
                                                   make_double3(s.g.dx, s.g.dy, s.g.dz), D.partial.get(),
                                                   D.maxima.get());
  kFinish<<<2, kRedThreads, 0, s.stream>>>(D.partial.get(), kRedBlocks, D.out.get());
  launchCheck("gauss");
  double sums[2];
  unsigned long long m[2];
  hipCheck(hipMemcpyAsync(sums, D.out.get(), sizeof(sums), hipMemcpyDeviceToHost, s.stream), "hipMemcpyAsync");
  hipCheck(hipMemcpyAsync(m, D.maxima.get(), sizeof(m), hipMemcpyDeviceToHost, s.stream), "hipMemcpyAsync");
  hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
  out[0] = sums[0];
  out[1] = sums[1];
  std::memcpy(&out[2], &m[0], sizeof(double));
  std::memcpy(&out[3], &m[1], sizeof(double));
}

// ---------------------------------------------------------------------------
// Moving-window energy ledger.

namespace {

// Ledger accumulator slots (compensated pairs).
constexpr int kLedgerOverwrite = 0, kLedgerInserted = 1, kLedgerCropped = 2, kLedgerRemap = 3, kLedgerFace = 4;
constexpr int kLedgerSlots = 10, kLedgerWork = 32;
constexpr int kLedgerFaceBlocks = 64;

__device__ void ledgerAdd(double* acc, int slot, double v) {
  Acc a{acc[2 * slot], acc[2 * slot + 1]};
  a.add(v);
  acc[2 * slot] = a.s;
  acc[2 * slot + 1] = a.c;
}

// op 0: begin (overwrite delta := 0); op 1: inlet overwrite measured
// (moving_window_fields.cpp:283-287); op 2: trapezoid face flux of the
// finished step (430-435); op 3: the shift's removed and inserted planes
// (503-509).
__global__ void kLedgerUpdate(double* acc, double* w, int op, double halfDt) {
  if (threadIdx.x || blockIdx.x) return;
  if (op == 0) {
    for (int f = 0; f < 6; ++f) w[6 + f] = 0.0;
  } else if (op == 1) {
    ledgerAdd(acc, kLedgerOverwrite, w[25] - w[24]);
    // the inlet-only power has no upper z face
    for (int f = 0; f < 6; ++f) w[6 + f] = f == 5 ? 0.0 : w[18 + f] - w[12 + f];
  } else if (op == 2) {
    for (int f = 0; f < 6; ++f) ledgerAdd(acc, kLedgerFace + f, halfDt * (w[f] + w[18 + f] - w[6 + f]));
  } else {
    ledgerAdd(acc, kLedgerRemap, w[27] - w[26]);
    ledgerAdd(acc, kLedgerInserted, w[27]);
    ledgerAdd(acc, kLedgerCropped, w[26]);
  }
}

template <class R>
__global__ void kInletOverwrite(R* ex, R* ey, const R* inletX, const R* inletY, int nx, int ny) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int planeX = nx * (ny + 1), planeY = (nx + 1) * ny;
  if (inletX && t < planeX) ex[t] = inletX[t];
  if (inletY && t < planeY) ey[t] = inletY[t];
}

// planeEnergy of moving_window_fields.cpp:387-397 over cells [lo, hi) into
// a device slot.
template <class R>
void energyTo(FieldStateT<R>& s, int3 lo, int3 hi, double* dst) {
  DiagnosticScratch& D = s.diag;
  ensure(D.partial, 2 * 6 * kRedBlocks);
  kEnergyBox<<<kRedBlocks, kRedThreads, 0, s.stream>>>(s.ptrs(), lo, hi, s.g.dx * s.g.dy * s.g.dz,
                                                          D.partial.get());
  kFinish<<<1, kRedThreads, 0, s.stream>>>(D.partial.get(), kRedBlocks, dst);
  launchCheck("ledger energy");
}

// boundaryPower of moving_window_fields.cpp:398-419 over the cell box
// [lo, hi) into six device slots.
template <class R>
void powerTo(FieldStateT<R>& s, int3 lo, int3 hi, double* dst) {
  DiagnosticScratch& D = s.diag;
  ensure(D.partial, 2 * 6 * kRedBlocks);
  const double3 area = make_double3(s.g.dy * s.g.dz, s.g.dx * s.g.dz, s.g.dx * s.g.dy);
  kFacePower<<<dim3(kLedgerFaceBlocks, 6), kRedThreads, 0, s.stream>>>(s.ptrs(), lo, hi, area, D.partial.get());
  kFinish<<<6, kRedThreads, 0, s.stream>>>(D.partial.get(), kLedgerFaceBlocks, dst);
  launchCheck("ledger power");
}

// The physical aperture (transverse CPML excluded) at z cells [k0, k1).
template <class R>
void apertureBox(const FieldStateT<R>& s, int k0, int k1, int3& lo, int3& hi) {
  lo = make_int3(static_cast<int>(s.cpml.cells[0]), static_cast<int>(s.cpml.cells[1]), k0);
  hi = make_int3(static_cast<int>(s.g.nx - s.cpml.cells[0]), static_cast<int>(s.g.ny - s.cpml.cells[1]), k1);
}

template <class R>
void ledgerOp(FieldStateT<R>& s, int op) {
  kLedgerUpdate<<<1, 1, 0, s.stream>>>(s.ledger.acc.get(), s.ledger.work.get(), op, 0.5 * s.g.dt);
  launchCheck("ledger update");
}

}  // namespace

template <class R>
void ledgerEnable(FieldStateT<R>& s, bool enable) {
  if (!s.window) throw std::logic_error("No moving window is installed");
  // MovingWindowFields::enableEnergyLedger, moving_window_fields.cpp:541-545
  if (s.established) throw std::logic_error("Moving-window energy ledger must be configured before the first step");
  EnergyLedgerDev& L = s.ledger;
  if (enable && !L.enabled) {
    L.acc.resize(2 * kLedgerSlots);
    L.work.resize(kLedgerWork);
    hipCheck(hipMemsetAsync(L.acc.get(), 0, 2 * kLedgerSlots * sizeof(double), s.stream), "hipMemsetAsync");
    hipCheck(hipMemsetAsync(L.work.get(), 0, kLedgerWork * sizeof(double), s.stream), "hipMemsetAsync");
    L.overwriteEvents = L.fluxSteps = L.remapSteps = 0;
    L.stepFinished = true;
  }
  L.enabled = enable;
}

template <class R>
void ledgerBegin(FieldStateT<R>& s) {
  if (!s.ledger.enabled) return;
  int3 lo, hi;
  apertureBox(s, 0, static_cast<int>(s.g.nz), lo, hi);
  powerTo(s, lo, hi, s.ledger.work.get());
  ledgerOp(s, 0);
  s.ledger.stepFinished = false;
}

template <class R>
void ledgerCommitElectric(FieldStateT<R>& s, const FieldPtrsT<R>& f) {
  R* const none[3] = {nullptr, nullptr, nullptr};
  launchElectric(s, f, 0, none, static_cast<const R*>(nullptr), static_cast<const R*>(nullptr));
  double* w = s.ledger.work.get();
  int3 lo, hi;
  apertureBox(s, 0, 1, lo, hi);
  energyTo(s, lo, hi, w + 24);
  powerTo(s, lo, hi, w + 12);
  const int nx = static_cast<int>(s.g.nx), ny = static_cast<int>(s.g.ny);
  kInletOverwrite<R><<<ceilDiv(static_cast<std::size_t>(nx + 1) * (ny + 1), 256), 256, 0, s.stream>>>(
      f.e[0], f.e[1], s.inletPlane[0], s.inletPlane[1], nx, ny);
  launchCheck("inlet overwrite");
  energyTo(s, lo, hi, w + 25);
  powerTo(s, lo, hi, w + 18);
  ledgerOp(s, 1);
  ++s.ledger.overwriteEvents;
}

template <class R>
void windowCommitElectric(FieldStateT<R>& s, const FieldPtrsT<R>& f) {
  if (s.ledger.enabled) {
    ledgerCommitElectric(s, f);
    return;
  }
  R* const none[3] = {nullptr, nullptr, nullptr};
  launchElectric(s, f, 0, none, s.inletPlane[0], s.inletPlane[1]);
}

template <class R>
void ledgerFinish(FieldStateT<R>& s) {
  if (!s.ledger.enabled) return;
  if (s.phase != 3 || s.ledger.stepFinished)
    throw std::logic_error("Moving-window energy flux requires one completed MPI field step");
  int3 lo, hi;
  apertureBox(s, 0, static_cast<int>(s.g.nz), lo, hi);
  powerTo(s, lo, hi, s.ledger.work.get() + 18);
  ledgerOp(s, 2);
  ++s.ledger.fluxSteps;
  s.ledger.stepFinished = true;
}

template <class R>
void ledgerBeforeShift(FieldStateT<R>& s) {
  if (!s.ledger.enabled) return;
  if (!s.ledger.stepFinished) throw std::logic_error("Moving-window energy flux must finish after MPI and before shift");
  int3 lo, hi;
  apertureBox(s, static_cast<int>(s.g.nz) - 1, static_cast<int>(s.g.nz), lo, hi);
  energyTo(s, lo, hi, s.ledger.work.get() + 26);
}

template <class R>
void ledgerAfterShift(FieldStateT<R>& s) {
  if (!s.ledger.enabled) return;
  int3 lo, hi;
  apertureBox(s, 0, 1, lo, hi);
  energyTo(s, lo, hi, s.ledger.work.get() + 27);
  ledgerOp(s, 3);
  ++s.ledger.remapSteps;
}

template <class R>
void ledgerRead(FieldStateT<R>& s, double out[13]) {
  if (!s.ledger.enabled || !s.ledger.acc.size()) throw std::logic_error("The moving-window energy ledger is not enabled");
  double acc[2 * kLedgerSlots];
  hipCheck(hipMemcpyAsync(acc, s.ledger.acc.get(), sizeof(acc), hipMemcpyDeviceToHost, s.stream), "hipMemcpyAsync");
  hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
  for (int k = 0; k < kLedgerSlots; ++k) out[k] = acc[2 * k] + acc[2 * k + 1];
  out[10] = static_cast<double>(s.ledger.overwriteEvents);
  out[11] = static_cast<double>(s.ledger.fluxSteps);
  out[12] = static_cast<double>(s.ledger.remapSteps);
}

// ---------------------------------------------------------------------------
// Particle sums of the energy ledger (Simulation, simulation.cpp:3591-3625).

namespace {

constexpr int kEnergyChannels = 9;

struct ParticleSumArgs {
  const double* x[3];  // plain form: absolute positions, u, mass
  const double* u[3];
  const double* m;
  const int* cell[3];  // FP32-split form: lattice cell + offset, u, mass
  const float* off[3];
  const float* uf[3];
  const float* mf;
  const int* state;
  int n, split;
  double h[3];
  double time;
  BoostDev frame;
};

__device__ __forceinline__ double gammaOfU(const double u[3]) {
  const double scale = fmax(1.0, fmax(fabs(u[0]), fmax(fabs(u[1]), fabs(u[2]))));
  const double inv = 1.0 / scale;
  const double a = u[0] * inv, b = u[1] * inv, c = u[2] * inv;
  return scale * sqrt(inv * inv + a * a + b * b + c * c);
}

// One entry's channels: rest u^2 / (gamma + 1), rest gamma, m / m_e, and
// the represented-weighted lab gamma, box gamma and lab z (simulation.cpp:
// 3593-3620), then ParticleEnergyMoment::sample's lab kinetic energy and
// c p_z in the box and the lab (particle_energy_audit.h:20-40).
__device__ void particleChannels(const ParticleSumArgs& A, int p, double out[kEnergyChannels]) {
  double u[3], z, m;
  if (A.split) {
    for (int a = 0; a < 3; ++a) u[a] = static_cast<double>(A.uf[a][p]);
    z = (static_cast<double>(A.cell[2][p]) + static_cast<double>(A.off[2][p])) * A.h[2];
    m = static_cast<double>(A.mf[p]);
  } else {
    for (int a = 0; a < 3; ++a) u[a] = A.u[a][p];
    z = A.x[2][p];
    m = A.m[p];
  }
  const BoostDev& F = A.frame;
  const double gammaBox = gammaOfU(u);
  const double rest = m * kC * kC;
  const double u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const double represented = m / kMe;
  // properVelocityBoxToLab through the positive light-front component
  // (boostframe.cpp:110-164)
  const double mt2 = fma(u[0], u[0], fma(u[1], u[1], 1.0));
  const double positive = (u[2] >= 0.0 ? gammaBox + u[2] : mt2 / (gammaBox - u[2])) * (F.gamma + F.gammaBeta);
  const double lab[3] = {u[0], u[1], 0.5 * (positive - mt2 / positive)};
  const double gammaLab = gammaOfU(lab);
  const double labZ = F.labZ0 + fma(F.gammaBeta * F.c0, A.time, F.gamma * (z - F.boxZ0));
  const double u2Lab = lab[0] * lab[0] + lab[1] * lab[1] + lab[2] * lab[2];
  out[0] = rest * u2 / (gammaBox + 1.0);
  out[1] = rest * gammaBox;
  out[2] = represented;
  out[3] = represented * gammaLab;
  out[4] = represented * gammaBox;
  out[5] = represented * labZ;
  out[6] = rest * u2Lab / (sqrt(1.0 + u2Lab) + 1.0);
  out[7] = rest * u[2];
  out[8] = rest * lab[2];
}

__global__ void __launch_bounds__(kRedThreads) kParticleEnergy(ParticleSumArgs A, double* partial) {
  __shared__ double ss[kRedThreads], sc[kRedThreads];
  Acc a[kEnergyChannels];
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < A.n; p += gridDim.x * blockDim.x) {
    const int st = A.state[p];
    if (st != kStateActive && st != kStateHeld) continue;  // particles_ holds active and held entries
    double v[kEnergyChannels];
    particleChannels(A, p, v);
    for (int c = 0; c < kEnergyChannels; ++c) a[c].add(v[c]);
  }
  for (int c = 0; c < kEnergyChannels; ++c) {
    const Acc b = blockReduce(a[c], ss, sc);
    if (threadIdx.x == 0) {
      partial[2 * (c * gridDim.x + blockIdx.x)] = b.s;
      partial[2 * (c * gridDim.x + blockIdx.x) + 1] = b.c;
    }
  }
}

}  // namespace

template <class R, class Store>
void particleEnergySums(FieldStateT<R>& s, const CouplingState& C, const Store& P, double time,
                        double out[kEnergyChannels]) {
  ParticleSumArgs A{};
  A.n = static_cast<int>(P.count);
  A.state = P.state.get();
  A.h[0] = s.g.dx;
  A.h[1] = s.g.dy;
  A.h[2] = s.g.dz;
  A.time = time;
  A.frame = C.frame;
  if constexpr (std::is_same_v<Store, SplitStore>) {
    A.split = 1;
    for (int a = 0; a < 3; ++a) {
      A.cell[a] = P.cell[a].get();
      A.off[a] = P.s[a].get();
      A.uf[a] = P.s[3 + a].get();
    }
    A.mf = P.s[7].get();
  } else {
    for (int a = 0; a < 3; ++a) {
      A.x[a] = P.s[a].get();
      A.u[a] = P.s[3 + a].get();
    }
    A.m = P.s[7].get();
  }
  DiagnosticScratch& D = s.diag;
  ensure(D.partial, 2 * kEnergyChannels * kRedBlocks);
  ensure(D.out, 16);
  kParticleEnergy<<<kRedBlocks, kRedThreads, 0, s.stream>>>(A, D.partial.get());
  kFinish<<<kEnergyChannels, kRedThreads, 0, s.stream>>>(D.partial.get(), kRedBlocks, D.out.get());
  launchCheck("particle energy");
  hipCheck(hipMemcpyAsync(out, D.out.get(), kEnergyChannels * sizeof(double), hipMemcpyDeviceToHost, s.stream),
           "hipMemcpyAsync");
  hipCheck(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
}

template void particleEnergySums(FieldState&, const CouplingState&, const ParticleStore&, double, double*);
template void particleEnergySums(FieldState32&, const CouplingState&, const SplitStore&, double, double*);

#define APRL_DIAG_INSTANTIATE(R)                                                                              \
  template void ledgerEnable(FieldStateT<R>&, bool);                                                          \
  template void ledgerBegin(FieldStateT<R>&);                                                                 \
  template void ledgerCommitElectric(FieldStateT<R>&, const FieldPtrsT<R>&);                                  \
  template void windowCommitElectric(FieldStateT<R>&, const FieldPtrsT<R>&);                                  \
  template void ledgerFinish(FieldStateT<R>&);                                                                \
  template void ledgerBeforeShift(FieldStateT<R>&);                                                           \
  template void ledgerAfterShift(FieldStateT<R>&);                                                            \
  template void ledgerRead(FieldStateT<R>&, double*);                                                         \
  template double interiorEnergy(FieldStateT<R>&, const std::size_t*, const std::size_t*);                    \
  template void facePower(FieldStateT<R>&, const std::size_t*, const std::size_t*, double*);                  \
  template void continuityResidual(FieldStateT<R>&, const double*, const double*, const RhoWorkspace&,         \
                                   double*);                                                                  \
  template RhoWorkspace diagnosticWorkspace(FieldStateT<R>&);                                                 \
  template RhoWorkspace noAllocWorkspace(FieldStateT<R>&, std::size_t, double**, double**);                   \
  template void gaussResidual(FieldStateT<R>&, const DeviceArray<double>&, const std::size_t*,                \
                              const std::size_t*, double*);
APRL_DIAG_INSTANTIATE(double)
APRL_DIAG_INSTANTIATE(float)
#undef APRL_DIAG_INSTANTIATE

}  // namespace aprl_gpu
