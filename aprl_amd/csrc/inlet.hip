// Charge-matched moving-window inlet (ChargeMatchedInlet,
// charge_matched_inlet.h:16-72) and the centered normal reference of the
// inlet Ez halo (moving_window_fields.cpp:225-234), for both forms.
//
// After a shift the new layer's Ez on the interior aperture is set so that
// the discrete Gauss law holds at the inlet vertices:
//   Ez(i,j,0) = Ez(i,j,1) + dz (div_xy E(i,j,1) - rho(i,j) / eps0),
// with rho the shaped charge at the inlet plane of the particles whose
// longitudinal shape reaches node k = 1. The charge is summed in 64-bit
// fixed point with one exponent from a device-side bound (largest |q w_z|
// times the contributing count), so the sums do not depend on the order of
// the atomics; the plane update runs in FP64 and rounds once on store.
#include <cmath>

#include "charge_basis.hpp"
#include "inlet.hpp"

namespace aprl_gpu {

namespace {

struct InletArgs {
  const double* x[3];  // absolute positions (plain form), or
  const int* cell[3];  // lattice cell + offset (split form)
  const float* off[3];
  int oc[3];
  const double* qd;
  const float* qf;
  const int* state;
  int n;
  double o[3], h[3];
  int nx, ny, shape, xBinomial;
  double invVolume;
  unsigned long long* words;  // [0] largest |q w_z| (double bits), [1] contributing particles
  unsigned long long* acc;    // (nx + 1)(ny + 1) fixed-point charge
};

// Node-1 weight and transverse position of one particle, or false when it
// does not reach the inlet (ChargeMatchedInlet::particle's filters).
__device__ bool inletWeights(const InletArgs& A, int p, double& wz, double& x, double& y, double& q) {
  // particles_ with the held ones (pending particles near the inlet are
  // held entries of the store here); carriers and removed entries do not count
  if (A.state[p] != kStateActive && A.state[p] != kStateHeld) return false;
  double l[3];
  for (int a = 0; a < 3; ++a)
    l[a] = A.x[0] ? (A.x[a][p] - A.o[a]) / A.h[a]
                  : static_cast<double>(A.cell[a][p] - A.oc[a]) + static_cast<double>(A.off[a][p]);
  double v[12];
  const int k0 = chargeBasis(A.shape, l[2], v);
  wz = 0.0;
#pragma unroll
  for (int c = 0; c < 12; ++c)
    if (k0 + c == 1) wz = v[c];
  if (wz == 0.0) return false;
  x = l[0];
  y = l[1];
  q = A.qd ? A.qd[p] : static_cast<double>(A.qf[p]);
  if (y < 0. || y > A.ny) return false;
  if (A.xBinomial ? (x <= -2. || x >= A.nx + 2.) : (x < 0. || x > A.nx)) return false;
  return true;
}

__global__ void kInletBound(InletArgs A) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= A.n) return;
  double wz, x, y, q;
  if (!inletWeights(A, p, wz, x, y, q)) return;
  const double m = fabs(q * wz);
  atomicMax(A.words, static_cast<unsigned long long>(__double_as_longlong(m)));
  atomicAdd(A.words + 1, 1ull);
}

__device__ int inletShift(const unsigned long long* words, double invVolume) {
  const double bound = __longlong_as_double(static_cast<long long>(words[0])) * static_cast<double>(words[1]) *
                       invVolume;
  if (!(bound > 0.0)) return 0;
  int e2 = 0;
  (void)frexp(bound, &e2);
  return 61 - e2;  // every partial sum stays below 2^61
}

__global__ void kInletCharge(InletArgs A) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= A.n) return;
  double wz, x, y, q;
  if (!inletWeights(A, p, wz, x, y, q)) return;
  const int shift = inletShift(A.words, A.invVolume);
  const long long ix = static_cast<long long>(floor(x));
  const long long j = static_cast<long long>(floor(y));
  const double fx = x - ix, fy = y - j, scale = q * A.invVolume;
  double wx[4];
  if (A.xBinomial) {
    wx[0] = 0.25 * (1. - fx); wx[1] = 0.5 - 0.25 * fx; wx[2] = 0.25 + 0.25 * fx; wx[3] = 0.25 * fx;
  } else {
    wx[0] = 1. - fx; wx[1] = fx; wx[2] = 0.; wx[3] = 0.;
  }
  const int nX = A.xBinomial ? 4 : 2;
  const long long first = ix - (A.xBinomial ? 1 : 0);
#pragma unroll
  for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      if (a >= nX) continue;
      const long long i = first + a;
      if (i < 0 || i > A.nx || j + b > A.ny) continue;
      const double value = scale * wx[a] * (b ? fy : 1. - fy) * wz;
      atomicAdd(A.acc + (j + b) * (A.nx + 1) + i, static_cast<unsigned long long>(llrint(ldexp(value, shift))));
    }
}

template <class R>
__global__ void kInletComplete(const R* ex1, const R* ey1, const R* ez1, R* ez0, unsigned long long* acc,
                               const unsigned long long* words, int nx, int ny, int gx, int gy, double dx, double dy,
                               double dz, double invVolume) {
  const int shift = inletShift(words, invVolume);
  const int np = (nx + 1) * (ny + 1);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < np; t += gridDim.x * blockDim.x) {
    const int i = t % (nx + 1), j = t / (nx + 1);
    const double rho = ldexp(static_cast<double>(static_cast<long long>(acc[t])), -shift);
    acc[t] = 0ull;
    if (j < gy + 1 || j >= ny - gy || i < gx + 1 || i >= nx - gx) continue;
    // Ex is nx wide with ny + 1 rows, Ey nx + 1 wide with ny rows (k = 1 planes)
    const double divxy = (static_cast<double>(ex1[j * nx + i]) - static_cast<double>(ex1[j * nx + i - 1])) / dx +
                         (static_cast<double>(ey1[j * (nx + 1) + i]) - static_cast<double>(ey1[(j - 1) * (nx + 1) + i])) / dy;
    const double value = static_cast<double>(ez1[t]) + dz * (divxy - rho / kEps0);
    ez0[t] = static_cast<R>(value);
  }
}

// lower Ez halo -= Ez(k=0) - reference Ez at the offset, on the aperture.
template <class R>
__global__ void kCenterInletHalo(R* halo, const R* ez0, const R* ref, int nx, int ny, int gx, int gy) {
  const int np = (nx + 1) * (ny + 1);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < np; t += gridDim.x * blockDim.x) {
    const int i = t % (nx + 1), j = t / (nx + 1);
    if (j < gy + 1 || j >= ny - gy || i < gx + 1 || i >= nx - gx) continue;
    halo[t] = static_cast<R>(static_cast<double>(halo[t]) -
                             (static_cast<double>(ez0[t]) - static_cast<double>(ref[t])));
  }
}

template <class R>
void completeInlet(FieldStateT<R>& F, InletArgs A, int shape, bool xBinomial, hipStream_t st) {
  if (!F.window || !F.chargeMatch) throw std::logic_error("The window has no charge-matched inlet");
  const aprl_grid_geometry& g = F.g;
  const std::size_t gx = F.cpml.cells[0], gy = F.cpml.cells[1];
  if (g.nz < 2 || gx + 1 >= g.nx - gx || gy + 1 >= g.ny - gy)
    throw std::invalid_argument("Charge-matched inlet needs an interior physical aperture");
  const std::size_t np = (g.nx + 1) * (g.ny + 1);
  if (F.inletAcc.size() != np) {
    F.inletAcc.resize(np);
    hipCheck(hipMemsetAsync(F.inletAcc.get(), 0, np * sizeof(unsigned long long), st), "hipMemsetAsync");
  }
  if (!F.inletWords.size()) F.inletWords.resize(2);
  hipCheck(hipMemsetAsync(F.inletWords.get(), 0, 2 * sizeof(unsigned long long), st), "hipMemsetAsync");
  A.h[0] = g.dx; A.h[1] = g.dy; A.h[2] = g.dz;
  A.nx = static_cast<int>(g.nx);
  A.ny = static_cast<int>(g.ny);
  A.shape = shape;
  A.xBinomial = xBinomial ? 1 : 0;
  A.invVolume = 1. / (g.dx * g.dy * g.dz);
  A.words = F.inletWords.get();
  A.acc = F.inletAcc.get();
  if (A.n > 0) {
    kInletBound<<<ceilDiv(A.n, 256), 256, 0, st>>>(A);
    kInletCharge<<<ceilDiv(A.n, 256), 256, 0, st>>>(A);
  }
  // k = 1 planes of Ex, Ey, Ez and the new layer of Ez
  const R* ex1 = F.base(0) + F.plane(0);
  const R* ey1 = F.base(1) + F.plane(1);
  const R* ez1 = F.base(2) + F.plane(2);
  kInletComplete<R><<<ceilDiv(static_cast<int>(np), 256), 256, 0, st>>>(
      ex1, ey1, ez1, F.base(2), F.inletAcc.get(), F.inletWords.get(), A.nx, A.ny, static_cast<int>(gx),
      static_cast<int>(gy), g.dx, g.dy, g.dz, A.invVolume);
  launchCheck("charge-matched inlet");
}

}  // namespace

void completeInletPlain(FieldState& F, const ParticleStore& P, const double origin[3], int shape, bool xBinomial) {
  InletArgs A{};
  for (int a = 0; a < 3; ++a) {
    A.x[a] = P.s[a].get();
    A.o[a] = origin[a];
  }
  A.qd = P.s[6].get();
  A.state = P.state.get();
  A.n = static_cast<int>(P.count);
  completeInlet(F, A, shape, xBinomial, F.stream);
}

void completeInletSplit(FieldState32& F, const SplitStore& P, const double origin[3], int shape, bool xBinomial) {
  InletArgs A{};
  const double h[3] = {F.g.dx, F.g.dy, F.g.dz};
  for (int a = 0; a < 3; ++a) {
    A.cell[a] = P.cell[a].get();
    A.off[a] = P.s[a].get();
    const double r = origin[a] / h[a];
    if (!std::isfinite(r) || std::fabs(r) >= 1.0e9 || std::fabs(r - std::nearbyint(r)) > 1e-6)
      throw std::invalid_argument("An FP32-split grid origin must lie on the cell lattice (origin = integer * spacing)");
    A.oc[a] = static_cast<int>(std::nearbyint(r));
  }
  A.qf = P.s[6].get();
  A.state = P.state.get();
  A.n = static_cast<int>(P.count);
  completeInlet(F, A, shape, xBinomial, F.stream);
}

template <class R>
void centerInletHalo(FieldStateT<R>& F, R* halo, const R* ref) {
  const aprl_grid_geometry& g = F.g;
  const int np = static_cast<int>((g.nx + 1) * (g.ny + 1));
  kCenterInletHalo<R><<<ceilDiv(np, 256), 256, 0, F.stream>>>(halo, F.base(2), ref, static_cast<int>(g.nx),
                                                                static_cast<int>(g.ny),
                                                                static_cast<int>(F.cpml.cells[0]),
                                                                static_cast<int>(F.cpml.cells[1]));
  launchCheck("centered inlet halo");
}
template void centerInletHalo<double>(FieldStateT<double>&, double*, const double*);
template void centerInletHalo<float>(FieldStateT<float>&, float*, const float*);

}  // namespace aprl_gpu
