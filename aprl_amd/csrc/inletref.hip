// The moving window's inlet reference on the device: BlockedCowanReference
// (blocked_cowan_reference.cpp) for one rank, held in device memory instead
// of block files, and movingWindowFieldValue (simulation_moving_window.cpp:
// 280-294) evaluated into the window's planes by a kernel.
//
// Preparation, once per run:
//   1. configure / validate (315-384) and the sine modes of the 27-point
//      Cowan stencil with the translation term (makeModes, 505-551) on the
//      host, with the reference's checks and messages;
//   2. the source charge on the interior vertices (deposit 595-655, or the
//      production longitudinal basis with linear / x-binomial3 transverse
//      weights, depositQuintic 559-593): every contribution formed in the
//      reference's arithmetic, summed in 64-bit fixed point (a bound pass
//      fixes one exponent), so the sum does not depend on the order;
//   3. rho / eps0, the 2-D DST-I (FFTW's RODFT00, y_k = 2 sum_j x_j
//      sin(pi (j+1)(k+1) / (n+1)), along x then y; one workgroup per line,
//      the line in LDS, the sine table in HBM), one Thomas sweep per mode
//      along z (a lane per mode, coalesced over the modes), the inverse DST
//      scaled by 1 / (4 (nx+1)(ny+1)) (solve, 709-770);
//   4. the native E(0), B(-dt/2) from the potential (ensureField, 822-950).
// Per step, referencePlane() fills a plane of the window from the prepared
// field in one kernel: no host callback and no upload.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#include "handles.hpp"
#include "inletref.hpp"
#include "charge_basis.hpp"

namespace aprl_gpu {
namespace {

constexpr int kThreads = 256;
constexpr double kEps = std::numeric_limits<double>::epsilon();

enum : unsigned {
  kErrOutside = 1u << 0,   // CIC: particle outside the global E/B domain (out_of_range)
  kErrNonfinite = 1u << 1, // nonfinite contribution
  kErrDirichlet = 1u << 2, // CIC support touches the Dirichlet surface
  kErrShapedWall = 1u << 3,// shaped support touches the Dirichlet surface
  kErrPivot = 1u << 4,     // nonpositive / nonfinite Thomas pivot
};

struct Lattice {
  long long nx, ny, nz;         // cells of the global grid
  long long lower[3], upper[3];  // Dirichlet planes (guards)
  long long nodes[3];            // interior vertices per axis
  double h[3], origin[3];
};

__device__ __forceinline__ bool interiorVertex(const Lattice& L, long long i, long long j, long long k) {
  return i > L.lower[0] && i < L.upper[0] && j > L.lower[1] && j < L.upper[1] && k > L.lower[2] && k < L.upper[2];
}
__device__ __forceinline__ long long interiorIndex(const Lattice& L, long long i, long long j, long long k) {
  return ((k - L.lower[2] - 1) * L.nodes[1] + (j - L.lower[1] - 1)) * L.nodes[0] + (i - L.lower[0] - 1);
}

struct Source {
  const double* x;  // interleaved positions
  const double* q;
  long long n;
  int shape;        // -1 CIC, 0 compensated quintic, 2 linear-binomial3
  int xBinomial;
  int absolute;     // 1: accumulate |value| (the bound pass)
  int shift;
};

__device__ __forceinline__ void addFixed(unsigned long long* acc, long long at, double value, const Source& S) {
  atomicAdd(acc + at, static_cast<unsigned long long>(llrint(ldexp(S.absolute ? fabs(value) : value, S.shift))));
}

// deposit (CIC), blocked_cowan_reference.cpp:600-629: the cell and
// fraction clamps and the eight corners in the reference's order.
__global__ void __launch_bounds__(kThreads) kRefCic(Lattice L, Source S, double inverseVolume, unsigned long long* acc,
                                                    unsigned* err) {
  const long long p = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x;
  if (p >= S.n) return;
  const long long cells[3] = {L.nx, L.ny, L.nz};
  long long cell[3];
  double fraction[3];
  unsigned e = 0;
  for (int a = 0; a < 3; ++a) {
    const double position = S.x[3 * p + a];
    const double relative = position - L.origin[a], extent = cells[a] * L.h[a];
    const double tolerance = 64. * kEps * fmax(1., fmax(fabs(L.origin[a]), extent));
    if (!isfinite(position) || relative < -tolerance || relative > extent + tolerance) {
      atomicOr(err, kErrOutside);
      return;
    }
    const double normalized = fmax(0., fmin(static_cast<double>(cells[a]), relative / L.h[a]));
    cell[a] = normalized <= 0. ? 0
                               : (normalized >= static_cast<double>(cells[a]) ? cells[a] - 1
                                                                              : static_cast<long long>(floor(normalized)));
    fraction[a] = fmax(0., fmin(1., normalized - static_cast<double>(cell[a])));
  }
  const double q = S.q[p];
  for (int c = 0; c < 2; ++c)
    for (int b = 0; b < 2; ++b)
      for (int a = 0; a < 2; ++a) {
        const double value = q * inverseVolume * (a ? fraction[0] : 1. - fraction[0]) *
                             (b ? fraction[1] : 1. - fraction[1]) * (c ? fraction[2] : 1. - fraction[2]);
        if (!isfinite(value)) {
          e |= kErrNonfinite;
          continue;
        }
        if (value == 0.) continue;
        const long long i = cell[0] + a, j = cell[1] + b, k = cell[2] + c;
        if (!interiorVertex(L, i, j, k)) {
          e |= kErrDirichlet;
          continue;
        }
        addFixed(acc, interiorIndex(L, i, j, k), value, S);
      }
  if (e) atomicOr(err, e);
}

// depositQuintic, blocked_cowan_reference.cpp:566-584: the longitudinal
// basis times linear (or x-binomial3) x and linear y weights.
__global__ void __launch_bounds__(kThreads) kRefShaped(Lattice L, Source S, unsigned long long* acc, unsigned* err) {
  const long long p = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x;
  if (p >= S.n) return;
  const double x = (S.x[3 * p] - L.origin[0]) / L.h[0], y = (S.x[3 * p + 1] - L.origin[1]) / L.h[1],
               z = (S.x[3 * p + 2] - L.origin[2]) / L.h[2];
  const long long ix = static_cast<long long>(floor(x)), iy = static_cast<long long>(floor(y));
  const double fx = x - ix, fy = y - iy, scale = S.q[p] / (L.h[0] * L.h[1] * L.h[2]);
  double wz[12];
  const long long first = chargeBasis(S.shape, z, wz);
  const int count = S.shape == 2 ? 4 : 12;
  double wx[4];
  if (S.xBinomial) {
    wx[0] = 0.25 * (1. - fx);
    wx[1] = 0.5 - 0.25 * fx;
    wx[2] = 0.25 + 0.25 * fx;
    wx[3] = 0.25 * fx;
  } else {
    wx[0] = 1. - fx;
    wx[1] = fx;
    wx[2] = wx[3] = 0.;
  }
  const long long firstX = ix - (S.xBinomial ? 1 : 0);
  const int nX = S.xBinomial ? 4 : 2;
  unsigned e = 0;
  for (int c = 0; c < count; ++c) {
    const long long k = first + c;
    if (wz[c] == 0.) continue;
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < nX; ++i) {
        const double value = scale * wx[i] * (j ? fy : 1. - fy) * wz[c];
        if (value == 0.) continue;
        const long long nodeX = firstX + i, nodeY = iy + j;
        if (k < 0 || nodeX < 0 || nodeY < 0 || !interiorVertex(L, nodeX, nodeY, k)) {
          e |= kErrShapedWall;
          continue;
        }
        if (!isfinite(value)) {
          e |= kErrNonfinite;
          continue;
        }
        addFixed(acc, interiorIndex(L, nodeX, nodeY, k), value, S);
      }
  }
  if (e) atomicOr(err, e);
}

// Largest entry of a non-negative fixed-point array (one atomic per block).
__global__ void __launch_bounds__(kThreads) kMaxFixed(const unsigned long long* acc, long long n,
                                                      unsigned long long* out) {
  __shared__ unsigned long long sh[kThreads];
  unsigned long long m = 0;
  for (long long t = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x; t < n;
       t += static_cast<long long>(gridDim.x) * kThreads)
    m = max(m, acc[t]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) sh[threadIdx.x] = max(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, sh[0]);
}

// rho from the fixed-point sums (clearing them) and rho / eps0 for the
// solve; also the deposited charge per block (sum of rho * volume, in a
// fixed order inside the block).
__global__ void __launch_bounds__(kThreads) kRefRho(unsigned long long* acc, long long n, int shift, double* rho,
                                                    double* rhs, double volume, double* blockCharge) {
  __shared__ double sh[kThreads];
  double s = 0.;
  for (long long t = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x; t < n;
       t += static_cast<long long>(gridDim.x) * kThreads) {
    const double v = ldexp(static_cast<double>(static_cast<long long>(acc[t])), -shift);
    acc[t] = 0ull;
    rho[t] = v;
    rhs[t] = v / kEps0;
    s += v * volume;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) blockCharge[blockIdx.x] = sh[0];
}

// 2-D DST-I pieces: along x, one workgroup per (row, plane) line with the
// line in LDS; along y, one workgroup per (16 columns, plane) with the
// columns in LDS. table[j * n + k] = sin(pi (j + 1)(k + 1) / (n + 1)).
constexpr int kDstCols = 16;
__global__ void __launch_bounds__(kThreads) kDstX(const double* in, double* out, const double* table, int nx, int ny,
                                                  double factor) {
  extern __shared__ double line[];
  const long long base = (static_cast<long long>(blockIdx.y) * ny + blockIdx.x) * nx;
  for (int i = threadIdx.x; i < nx; i += kThreads) line[i] = in[base + i];
  __syncthreads();
  for (int k = threadIdx.x; k < nx; k += kThreads) {
    double s = 0.;
    for (int j = 0; j < nx; ++j) s += line[j] * table[static_cast<long long>(j) * nx + k];
    out[base + k] = 2. * s * factor;
  }
}
__global__ void __launch_bounds__(kThreads) kDstY(const double* in, double* out, const double* table, int nx, int ny,
                                                  double factor) {
  extern __shared__ double cols[];  // [ny][kDstCols]
  const int x0 = blockIdx.x * kDstCols;
  const long long plane = static_cast<long long>(blockIdx.y) * ny * nx;
  for (int t = threadIdx.x; t < ny * kDstCols; t += kThreads) {
    const int j = t / kDstCols, c = t % kDstCols;
    cols[t] = x0 + c < nx ? in[plane + static_cast<long long>(j) * nx + x0 + c] : 0.;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < ny * kDstCols; t += kThreads) {
    const int k = t / kDstCols, c = t % kDstCols;
    if (x0 + c >= nx) continue;
    double s = 0.;
    for (int j = 0; j < ny; ++j) s += cols[j * kDstCols + c] * table[static_cast<long long>(j) * ny + k];
    out[plane + static_cast<long long>(k) * nx + x0 + c] = 2. * s * factor;
  }
}

// solve's Thomas sweep (blocked_cowan_reference.cpp:725-749), one lane per
// mode m over the nz interior planes; rhs is overwritten by the solution.
__global__ void __launch_bounds__(kThreads) kThomas(double* rhs, double* cprime, const double* diagonal,
                                                    const double* off, long long area, long long nzn,
                                                    unsigned long long* minPivot, unsigned* err) {
  const long long m = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x;
  if (m >= area) return;
  const double dg = diagonal[m], of = off[m];
  double cprev = 0., dprev = 0., lowest = INFINITY;
  bool bad = false;
  for (long long k = 0; k < nzn; ++k) {
    const double pivot = dg - of * cprev;
    if (!isfinite(pivot) || !(pivot > 0.)) bad = true;
    lowest = fmin(lowest, pivot);
    cprev = of / pivot;
    dprev = (rhs[k * area + m] - of * dprev) / pivot;
    cprime[k * area + m] = cprev;
    rhs[k * area + m] = dprev;
  }
  double next = 0.;
  for (long long k = nzn; k > 0; --k) {
    const long long q = (k - 1) * area + m;
    const double v = rhs[q] - cprime[q] * next;
    rhs[q] = v;
    next = v;
  }
  if (bad) atomicOr(err, kErrPivot);
  // pivots are positive when valid: their bit patterns order as the values
  atomicMin(minPivot, static_cast<unsigned long long>(__double_as_longlong(bad ? 0.0 : lowest)));
}

// The interior solution onto the full vertex lattice (zero elsewhere).
__global__ void __launch_bounds__(kThreads) kScatterPhi(Lattice L, const double* interior, double* phi) {
  const long long vx = L.nx + 1, vy = L.ny + 1, n = vx * vy * (L.nz + 1);
  for (long long t = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x; t < n;
       t += static_cast<long long>(gridDim.x) * kThreads) {
    const long long i = t % vx, j = (t / vx) % vy, k = t / (vx * vy);
    phi[t] = interiorVertex(L, i, j, k) ? interior[interiorIndex(L, i, j, k)] : 0.;
  }
}

struct Smooth {
  double alpha[3], beta[3], delta[3];  // delta[axis]: deltaYZ, deltaZX, deltaXY
  double translation;
};

// smooth (blocked_cowan_reference.cpp:822-841) on every interior vertex,
// zero elsewhere; three arrays on the vertex lattice.
__global__ void __launch_bounds__(kThreads) kRefSmooth(Lattice L, Smooth C, const double* phi, double* f0, double* f1,
                                                       double* f2) {
  const long long vx = L.nx + 1, vy = L.ny + 1, n = vx * vy * (L.nz + 1);
  const long long stride[3] = {1, vx, vx * vy};
  double* out[3] = {f0, f1, f2};
  for (long long t = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x; t < n;
       t += static_cast<long long>(gridDim.x) * kThreads) {
    const long long i = t % vx, j = (t / vx) % vy, k = t / (vx * vy);
    if (!interiorVertex(L, i, j, k)) {
      for (int a = 0; a < 3; ++a) out[a][t] = 0.;
      continue;
    }
    for (int axis = 0; axis < 3; ++axis) {
      const int ta = (axis + 1) % 3, tb = (axis + 2) % 3;
      double value = C.alpha[axis] * phi[t];
      for (int sign = -1; sign <= 1; sign += 2) {
        const long long qa = sign < 0 ? t - stride[ta] : t + stride[ta];
        const long long qb = sign < 0 ? t - stride[tb] : t + stride[tb];
        value += C.beta[ta] * phi[qa];
        value += C.beta[tb] * phi[qb];
      }
      for (int a = -1; a <= 1; a += 2)
        for (int b = -1; b <= 1; b += 2) {
          long long s = a < 0 ? t - stride[ta] : t + stride[ta];
          s = b < 0 ? s - stride[tb] : s + stride[tb];
          value += C.delta[axis] * phi[s];
        }
      out[axis][t] = value - (axis == 2 ? C.translation * phi[t] : 0.);
    }
  }
}

struct FieldArgs {
  long long nx[6], ny[6], nz[6];
  double spacing[3];
  double magneticScale, lowerWeight, upperWeight;
};

// ensureField's native components (blocked_cowan_reference.cpp:919-943).
__global__ void __launch_bounds__(kThreads) kRefFields(Lattice L, FieldArgs A, int comp, const double* f0,
                                                       const double* f1, const double* f2, double* out) {
  const long long vx = L.nx + 1, full = vx * (L.ny + 1);
  const long long stride[3] = {1, vx, full};
  const double* filtered[3] = {f0, f1, f2};
  const long long cx = A.nx[comp], cy = A.ny[comp], n = cx * cy * A.nz[comp];
  for (long long t = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x; t < n;
       t += static_cast<long long>(gridDim.x) * kThreads) {
    const long long i = t % cx, j = (t / cx) % cy, k = t / (cx * cy);
    const long long q = k * full + j * vx + i;
    double value;
    if (comp < 3) {
      value = -(filtered[comp][q + stride[comp]] - filtered[comp][q]) / A.spacing[comp];
    } else if (comp < 5) {
      const int axis = comp == 3 ? 1 : 0;
      const double e0 = -(filtered[axis][q + stride[axis]] - filtered[axis][q]) / A.spacing[axis];
      const double e1 = -(filtered[axis][q + full + stride[axis]] - filtered[axis][q + full]) / A.spacing[axis];
      value = (comp == 3 ? -A.magneticScale : A.magneticScale) * (A.lowerWeight * e0 + A.upperWeight * e1);
    } else {
      value = 0.;
    }
    out[t] = value;
  }
}

// movingWindowFieldValue over one x-y plane: planes lo and hi, weight f.
template <class R>
__global__ void __launch_bounds__(kThreads) kRefPlane(const double* field, long long planeSize, long long lo,
                                                      long long hi, double f, R* out) {
  const long long t = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x;
  if (t >= planeSize) return;
  const double lower = field[lo * planeSize + t];
  out[t] = static_cast<R>(f == 0. ? lower : (1. - f) * lower + f * field[hi * planeSize + t]);
}

unsigned gridFor(long long n) { return static_cast<unsigned>(std::max(1LL, std::min<long long>(4096, (n + kThreads - 1) / kThreads))); }

void raiseReferenceErrors(unsigned e) {
  if (!e) return;
  if (e & kErrOutside) throw std::out_of_range("Blocked source particle outside global E/B domain");
  if (e & kErrNonfinite) throw std::invalid_argument("Nonfinite blocked CIC contribution");
  if (e & kErrDirichlet)
    throw std::invalid_argument("Blocked source CIC support touches Dirichlet surface; padding unsupported");
  if (e & kErrShapedWall) throw std::invalid_argument("Reference charge support touches Dirichlet surface");
  if (e & kErrPivot) throw std::runtime_error("Blocked Thomas matrix has nonpositive/nonfinite pivot");
}

double deltaFor(const CowanCoef& c, unsigned axis) { return axis == 0 ? c.deltaYZ : (axis == 1 ? c.deltaZX : c.deltaXY); }

}  // namespace

template <class R>
void referencePlane(const aprl_inlet_reference& ref, unsigned a, std::ptrdiff_t k, std::size_t step,
                    double translationBeta, R* out, hipStream_t st) {
  // simulation_moving_window.cpp:283-293
  const double z = static_cast<double>(k) - translationBeta * step;
  const std::size_t count = ref.g.nz + ((a == 0 || a == 1 || a == 5) ? 1 : 0);
  if (a >= 6 || !std::isfinite(z) || z < 0. || z > count - 1)
    throw std::out_of_range("Moving reference query outside prepared native field");
  const auto lo = static_cast<std::size_t>(std::floor(z)), hi = std::min(lo + 1, count - 1);
  const double f = z - lo;
  const long long planeSize = static_cast<long long>(ref.nx[a] * ref.ny[a]);
  // the reference's preparation runs on its own stream
  if (ref.stream != st) hipCheck(hipStreamSynchronize(ref.stream), "hipStreamSynchronize");
  kRefPlane<R><<<ceilDiv(planeSize, kThreads), kThreads, 0, st>>>(ref.field[a].get(), planeSize,
                                                                   static_cast<long long>(lo),
                                                                   static_cast<long long>(hi), f, out);
  launchCheck("reference plane");
}
template void referencePlane<double>(const aprl_inlet_reference&, unsigned, std::ptrdiff_t, std::size_t, double,
                                     double*, hipStream_t);
template void referencePlane<float>(const aprl_inlet_reference&, unsigned, std::ptrdiff_t, std::size_t, double,
                                    float*, hipStream_t);

}  // namespace aprl_gpu

using namespace aprl_gpu;

extern "C" aprl_status aprl_inlet_reference_create(const aprl_grid_geometry* global, const double* global_origin,
                                                   size_t n, const double* position3, const double* charge,
                                                   double beta, const size_t* guards, int source_shape,
                                                   int transverse_x_binomial3, aprl_inlet_reference** out,
                                                   aprl_inlet_reference_report* report) {
  return guarded([&] {
    if (!global || !global_origin || !guards || !out) throw std::invalid_argument("Null inlet reference argument");
    if (n && (!position3 || !charge)) throw std::invalid_argument("Null inlet reference source arrays");
    if (source_shape != -1 && source_shape != APRL_SHAPE_QUINTIC_COMPENSATED &&
        source_shape != APRL_SHAPE_LINEAR_BINOMIAL3)
      throw std::invalid_argument("Inlet reference source shape must be CIC, quintic-compensated or linear-binomial3");
    if (transverse_x_binomial3 && source_shape != APRL_SHAPE_LINEAR_BINOMIAL3)
      throw std::invalid_argument("Transverse three-point reference requires longitudinal linear-binomial3");
    const aprl_grid_geometry& g = *global;
    // configure / validate, blocked_cowan_reference.cpp:349-384
    if (g.solver != 1 || !std::isfinite(beta) || std::abs(beta) >= 1.)
      throw std::invalid_argument("Invalid blocked Cowan reference configuration");
    const std::size_t cells[3] = {g.nx, g.ny, g.nz};
    const double h[3] = {g.dx, g.dy, g.dz};
    Lattice L{};
    L.nx = static_cast<long long>(g.nx);
    L.ny = static_cast<long long>(g.ny);
    L.nz = static_cast<long long>(g.nz);
    for (int a = 0; a < 3; ++a) {
      if (!std::isfinite(h[a]) || h[a] <= 0. || !std::isfinite(global_origin[a]) || cells[a] < 3 ||
          guards[a] > (cells[a] - 3) / 2)
        throw std::invalid_argument("Invalid blocked reference geometry or guard/padding");
      L.lower[a] = static_cast<long long>(guards[a]);
      L.upper[a] = static_cast<long long>(cells[a] - guards[a]);
      L.nodes[a] = L.upper[a] - L.lower[a] - 1;
      L.h[a] = h[a];
      L.origin[a] = global_origin[a];
    }
    if (!std::isfinite(g.dt) || g.dt <= 0. || !std::isfinite(g.dx * g.dy * g.dz) || g.dx * g.dy * g.dz <= 0.)
      throw std::invalid_argument("Invalid blocked reference time step or cell volume");
    const CowanCoef cw = cowanForZDispersion(g.dx, g.dy, g.dz);
    const double speed = beta * kC;
    const double translation = speed * speed * kEps0 * kMu0;
    const double upperWeight = 0.5 + 0.5 * beta * kC * g.dt / g.dz, lowerWeight = 1. - upperWeight;
    const double magneticScale = beta * kC * kEps0 * kMu0;
    if (!(upperWeight >= 0. && upperWeight <= 1.))
      throw std::invalid_argument("Blocked initial B half-step shift exceeds one cell");
    // makeModes, blocked_cowan_reference.cpp:505-551
    double stencil[3][3][3] = {};
    for (unsigned axis = 0; axis < 3; ++axis) {
      const unsigned ta = (axis + 1) % 3, tb = (axis + 2) % 3;
      const double delta = deltaFor(cw, axis);
      const double inverse = 1. / (h[axis] * h[axis]);
      for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b) {
          const double weight = a == 0 ? (b == 0 ? cw.alpha[axis] : cw.beta[tb]) : (b == 0 ? cw.beta[ta] : delta);
          int p[3] = {1, 1, 1};
          p[ta] += a;
          p[tb] += b;
          stencil[p[2]][p[1]][p[0]] += 2. * weight * inverse;
          --p[axis];
          stencil[p[2]][p[1]][p[0]] -= weight * inverse;
          p[axis] += 2;
          stencil[p[2]][p[1]][p[0]] -= weight * inverse;
        }
    }
    const double translationZ = translation / (g.dz * g.dz);
    stencil[1][1][1] -= 2. * translationZ;
    stencil[0][1][1] += translationZ;
    stencil[2][1][1] += translationZ;
    const long long area = L.nodes[0] * L.nodes[1], nzn = L.nodes[2];
    std::vector<double> diagonal(static_cast<std::size_t>(area)), off(static_cast<std::size_t>(area));
    double minimumEigen = std::numeric_limits<double>::infinity();
    const double pi = std::acos(-1.);
    for (long long j = 0; j < L.nodes[1]; ++j)
      for (long long i = 0; i < L.nodes[0]; ++i) {
        const double tx = pi * (i + 1) / (L.nodes[0] + 1), ty = pi * (j + 1) / (L.nodes[1] + 1);
        double symbol[3] = {};
        for (unsigned z = 0; z < 3; ++z)
          for (unsigned y = 0; y < 3; ++y)
            for (unsigned x = 0; x < 3; ++x)
              symbol[z] += stencil[z][y][x] * std::cos((static_cast<int>(x) - 1) * tx) *
                           std::cos((static_cast<int>(y) - 1) * ty);
        const double scale = std::max(std::abs(symbol[1]), std::max(std::abs(symbol[0]), std::abs(symbol[2])));
        if (!std::isfinite(scale) || scale == 0. || std::abs(symbol[0] - symbol[2]) > 128. * kEps * scale)
          throw std::runtime_error("Blocked Cowan modal matrix is not finite symmetric");
        const std::size_t m = static_cast<std::size_t>(j * L.nodes[0] + i);
        diagonal[m] = symbol[1];
        off[m] = 0.5 * (symbol[0] + symbol[2]);
        const double eigenvalue = diagonal[m] - 2. * std::abs(off[m]) * std::cos(pi / (L.nodes[2] + 1));
        if (!(eigenvalue > 0.) || !std::isfinite(eigenvalue))
          throw std::runtime_error("Blocked Cowan modal matrix is not positive definite");
        minimumEigen = std::min(minimumEigen, eigenvalue);
      }
    // the source charge
    long double sourceCharge = 0.L;
    double absCharge = 0.;
    for (std::size_t p = 0; p < n; ++p) {
      if (!std::isfinite(charge[p])) throw std::invalid_argument("Nonfinite blocked source charge");
      sourceCharge += charge[p];
      absCharge += std::abs(charge[p]);
    }
    auto ref = std::make_unique<aprl_inlet_reference>();
    ref->g = g;
    for (int a = 0; a < 3; ++a) ref->origin[a] = global_origin[a];
    const std::size_t nxs[6] = {g.nx, g.nx + 1, g.nx + 1, g.nx + 1, g.nx, g.nx};
    const std::size_t nys[6] = {g.ny + 1, g.ny, g.ny + 1, g.ny, g.ny + 1, g.ny};
    for (int a = 0; a < 6; ++a) {
      ref->nx[a] = nxs[a];
      ref->ny[a] = nys[a];
      ref->nz[a] = g.nz + ((a == 0 || a == 1 || a == 5) ? 1 : 0);
    }
    hipStream_t st = ref->ownStream.get();
    ref->stream = st;
    const long long interior = area * nzn;
    DeviceArray<double> x(std::max<std::size_t>(3 * n, 1)), q(std::max<std::size_t>(n, 1));
    // host inputs are copied synchronously (pageable memory), before any launch reads them
    if (n) {
      hipCheck(hipMemcpy(x.get(), position3, 3 * n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
      hipCheck(hipMemcpy(q.get(), charge, n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    }
    DeviceArray<unsigned long long> acc(static_cast<std::size_t>(interior)), words(2);
    hipCheck(hipMemsetAsync(acc.get(), 0, static_cast<std::size_t>(interior) * sizeof(unsigned long long), st),
             "hipMemsetAsync");
    hipCheck(hipMemsetAsync(words.get(), 0, 2 * sizeof(unsigned long long), st), "hipMemsetAsync");
    unsigned* err = reinterpret_cast<unsigned*>(words.get() + 1);
    Source S{x.get(), q.get(), static_cast<long long>(n), source_shape, transverse_x_binomial3 ? 1 : 0, 1, 0};
    const double volume = g.dx * g.dy * g.dz, inverseVolume = 1. / volume;
    // bound pass: |contributions| summed per vertex against a coarse
    // exponent from the total |charge| (a contribution is at most 4 |q| / V)
    int e2 = 0;
    (void)std::frexp(4. * absCharge / volume + std::numeric_limits<double>::min(), &e2);
    S.shift = 62 - e2;
    const unsigned pBlocks = ceilDiv(std::max<std::size_t>(n, 1), kThreads);
    auto deposit = [&] {
      if (!n) return;
      if (source_shape < 0) kRefCic<<<pBlocks, kThreads, 0, st>>>(L, S, inverseVolume, acc.get(), err);
      else kRefShaped<<<pBlocks, kThreads, 0, st>>>(L, S, acc.get(), err);
      launchCheck("reference deposit");
    };
    deposit();
    kMaxFixed<<<gridFor(interior), kThreads, 0, st>>>(acc.get(), interior, words.get());
    launchCheck("reference deposit bound");
    unsigned long long peak = 0;
    unsigned e = 0;
    hipCheck(hipMemcpyAsync(&peak, words.get(), sizeof peak, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    hipCheck(hipMemcpyAsync(&e, err, sizeof e, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
    raiseReferenceErrors(e);
    // signed pass with the exponent of twice the largest vertex bound
    const double peakValue = std::ldexp(static_cast<double>(peak), -S.shift);
    (void)std::frexp(2. * peakValue + std::numeric_limits<double>::min(), &e2);
    S.shift = 62 - e2;
    S.absolute = 0;
    hipCheck(hipMemsetAsync(acc.get(), 0, static_cast<std::size_t>(interior) * sizeof(unsigned long long), st),
             "hipMemsetAsync");
    deposit();
    DeviceArray<double> rho(static_cast<std::size_t>(interior)), rhs(static_cast<std::size_t>(interior)),
        work(static_cast<std::size_t>(interior)), blockCharge(4096);
    const unsigned rBlocks = gridFor(interior);
    kRefRho<<<rBlocks, kThreads, 0, st>>>(acc.get(), interior, S.shift, rho.get(), rhs.get(), volume,
                                          blockCharge.get());
    launchCheck("reference density");
    acc.release();
    // sine tables and the forward 2-D DST of every interior plane
    const int mx = static_cast<int>(L.nodes[0]), my = static_cast<int>(L.nodes[1]);
    std::vector<double> tx(static_cast<std::size_t>(mx) * mx), ty(static_cast<std::size_t>(my) * my);
    for (int j = 0; j < mx; ++j)
      for (int k = 0; k < mx; ++k) tx[static_cast<std::size_t>(j) * mx + k] = std::sin(pi * double((j + 1) * (k + 1)) / double(mx + 1));
    for (int j = 0; j < my; ++j)
      for (int k = 0; k < my; ++k) ty[static_cast<std::size_t>(j) * my + k] = std::sin(pi * double((j + 1) * (k + 1)) / double(my + 1));
    DeviceArray<double> dtx(tx.size()), dty(ty.size()), dDiag(static_cast<std::size_t>(area)),
        dOff(static_cast<std::size_t>(area));
    hipCheck(hipMemcpy(dtx.get(), tx.data(), tx.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    hipCheck(hipMemcpy(dty.get(), ty.data(), ty.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    hipCheck(hipMemcpy(dDiag.get(), diagonal.data(), diagonal.size() * sizeof(double), hipMemcpyHostToDevice),
             "hipMemcpy");
    hipCheck(hipMemcpy(dOff.get(), off.data(), off.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    if (static_cast<std::size_t>(my) * kDstCols * sizeof(double) > 64 * 1024 ||
        static_cast<std::size_t>(mx) * sizeof(double) > 64 * 1024 || nzn > 65535)
      throw std::invalid_argument("Inlet reference grid exceeds the transform's launch limits (512 transverse, 65535 "
                                  "longitudinal interior nodes)");
    auto dst2 = [&](double* a, double* b, double factor) {
      kDstX<<<dim3(my, static_cast<unsigned>(nzn)), kThreads, mx * sizeof(double), st>>>(a, b, dtx.get(), mx, my, 1.);
      launchCheck("reference DST x");
      kDstY<<<dim3(ceilDiv(mx, kDstCols), static_cast<unsigned>(nzn)), kThreads, my * kDstCols * sizeof(double), st>>>(
          b, a, dty.get(), mx, my, factor);
      launchCheck("reference DST y");
    };
    dst2(rhs.get(), work.get(), 1.);
    // Thomas along z for every mode
    hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
    const unsigned long long pivotWords[2] = {0x7ff0000000000000ull, 0ull};  // +inf bits, cleared error word
    hipCheck(hipMemcpy(words.get(), pivotWords, sizeof pivotWords, hipMemcpyHostToDevice), "hipMemcpy");
    err = reinterpret_cast<unsigned*>(words.get() + 1);
    kThomas<<<ceilDiv(area, kThreads), kThreads, 0, st>>>(rhs.get(), work.get(), dDiag.get(), dOff.get(), area, nzn,
                                                          words.get(), err);
    launchCheck("reference Thomas");
    dst2(rhs.get(), work.get(), 1. / (4. * (mx + 1) * (my + 1)));
    unsigned long long pivotBits = 0;
    hipCheck(hipMemcpyAsync(&pivotBits, words.get(), sizeof pivotBits, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    hipCheck(hipMemcpyAsync(&e, err, sizeof e, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    std::vector<double> charges(rBlocks);
    hipCheck(hipMemcpyAsync(charges.data(), blockCharge.get(), rBlocks * sizeof(double), hipMemcpyDeviceToHost, st),
             "hipMemcpyAsync");
    hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
    raiseReferenceErrors(e);
    // the potential on the vertex lattice, its smoothing and the fields
    const long long vertices = (L.nx + 1) * (L.ny + 1) * (L.nz + 1);
    DeviceArray<double> phi(static_cast<std::size_t>(vertices));
    kScatterPhi<<<gridFor(vertices), kThreads, 0, st>>>(L, rhs.get(), phi.get());
    launchCheck("reference potential");
    rhs.release();
    work.release();
    DeviceArray<double> filtered[3];
    for (auto& f : filtered) f.resize(static_cast<std::size_t>(vertices));
    Smooth C{};
    for (int a = 0; a < 3; ++a) {
      C.alpha[a] = cw.alpha[a];
      C.beta[a] = cw.beta[a];
      C.delta[a] = deltaFor(cw, static_cast<unsigned>(a));
    }
    C.translation = translation;
    kRefSmooth<<<gridFor(vertices), kThreads, 0, st>>>(L, C, phi.get(), filtered[0].get(), filtered[1].get(),
                                                       filtered[2].get());
    launchCheck("reference smoothing");
    FieldArgs A{};
    for (int a = 0; a < 6; ++a) {
      A.nx[a] = static_cast<long long>(ref->nx[a]);
      A.ny[a] = static_cast<long long>(ref->ny[a]);
      A.nz[a] = static_cast<long long>(ref->nz[a]);
    }
    for (int a = 0; a < 3; ++a) A.spacing[a] = h[a];
    A.magneticScale = magneticScale;
    A.lowerWeight = lowerWeight;
    A.upperWeight = upperWeight;
    for (int c = 0; c < 6; ++c) {
      const long long count = A.nx[c] * A.ny[c] * A.nz[c];
      ref->field[c].resize(static_cast<std::size_t>(count));
      kRefFields<<<gridFor(count), kThreads, 0, st>>>(L, A, c, filtered[0].get(), filtered[1].get(),
                                                      filtered[2].get(), ref->field[c].get());
      launchCheck("reference fields");
    }
    hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (report) {
      long double deposited = 0.L;
      for (double v : charges) deposited += v;
      report->minimum_mode_eigenvalue = minimumEigen;
      double pivot;
      std::memcpy(&pivot, &pivotBits, sizeof pivot);
      report->minimum_pivot = pivot;
      report->source_charge = static_cast<double>(sourceCharge);
      report->deposited_charge = static_cast<double>(deposited);
      for (int a = 0; a < 3; ++a) report->interior_nodes[a] = static_cast<uint64_t>(L.nodes[a]);
    }
    *out = ref.release();
  });
}

extern "C" void aprl_inlet_reference_destroy(aprl_inlet_reference* reference) { delete reference; }

extern "C" aprl_status aprl_inlet_reference_download(aprl_inlet_reference* reference, int component, double* host,
                                                     size_t count) {
  return guarded([&] {
    if (!reference || !host) throw std::invalid_argument("Null inlet reference or output");
    if (component < 0 || component >= 6) throw std::invalid_argument("Inlet reference component must be 0..5");
    if (count != reference->field[component].size())
      throw std::invalid_argument("Inlet reference component size mismatch");
    hipCheck(hipStreamSynchronize(reference->stream), "hipStreamSynchronize");
    hipCheck(hipMemcpy(host, reference->field[component].get(), count * sizeof(double), hipMemcpyDeviceToHost),
             "hipMemcpy");
  });
}

extern "C" aprl_status aprl_inlet_reference_plane(aprl_inlet_reference* reference, int component, ptrdiff_t k,
                                                  size_t step, double translation_beta, double* host, size_t count) {
  return guarded([&] {
    if (!reference || !host) throw std::invalid_argument("Null inlet reference or output");
    if (component < 0 || component >= 6) throw std::invalid_argument("Inlet reference component must be 0..5");
    const unsigned a = static_cast<unsigned>(component);
    if (count != reference->nx[a] * reference->ny[a]) throw std::invalid_argument("Inlet reference plane size mismatch");
    DeviceArray<double> plane(count);
    referencePlane<double>(*reference, a, k, step, translation_beta, plane.get(), reference->stream);
    hipCheck(hipMemcpyAsync(host, plane.get(), count * sizeof(double), hipMemcpyDeviceToHost, reference->stream),
             "hipMemcpyAsync");
    hipCheck(hipStreamSynchronize(reference->stream), "hipStreamSynchronize");
  });
}
