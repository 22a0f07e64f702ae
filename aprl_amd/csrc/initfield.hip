// Gauss-consistent initial self-field on the device:
// EBGaussFieldInitializer::initialize (eb_initial_field.cpp:829-1095) for
// one rank. The particles' CIC charge (ChargeConservingCurrentDepositor::
// depositCharge, eb_deposition.cpp:168-193) drives a matrix-free conjugate
// gradient solve of the relativistic-Poisson 7-point or the Cowan 27-point
// operator on the vertex lattice with zero Dirichlet planes at the static
// field guard; E is the (Cowan-smoothed) gradient of the potential and B the
// rigid-translation field at -dt/2; the discrete Gauss residual is checked
// and the initial field energies reported.
//
// Sums the reference keeps in long double (mean velocity, moments, charge,
// CG dot products, Gauss and energy sums) are double-double here, each block
// summing in a fixed order and the blocks reduced in a fixed order, so the
// solve is reproducible; the CG step sizes are formed from them in long
// double on the host. The stencils, updates and field constructions use the
// reference's arithmetic order. The fields are built and checked in FP64 and
// then stored in the grid's precision.
#include <cmath>
#include <cstring>
#include <limits>
#include <sstream>

#include "ddsum.hpp"
#include "handles.hpp"
#include "particle_device.hpp"

namespace aprl_gpu {
namespace {

constexpr double kEpsilon = std::numeric_limits<double>::epsilon();
constexpr int kThreads = 256, kBlocks = 1024;

using namespace ddsum;

__device__ __forceinline__ void atomicMaxAbs(unsigned long long* slot, double v) {
  atomicMax(slot, static_cast<unsigned long long>(__double_as_longlong(fabs(v))));
}

// One store entry's position (m), proper velocity, charge and mass, and
// whether it takes part (active and held entries).
struct Source {
  const double* x[3];  // plain form: absolute positions, or
  const int* cell[3];  // FP32-split form: lattice cell + offset
  const float* off[3];
  double h[3];
  const double* ud[3];
  const float* uf[3];
  const double* qd;
  const double* md;
  const float* qf;
  const float* mf;
  const int* state;
  int n;
};
__device__ __forceinline__ bool sourceAt(const Source& S, int p, double x[3], double u[3], double& q, double& m) {
  const int st = S.state[p];
  if (st != kStateActive && st != kStateHeld) return false;
  for (int a = 0; a < 3; ++a) {
    x[a] = S.x[0] ? S.x[a][p] : (static_cast<double>(S.cell[a][p]) + static_cast<double>(S.off[a][p])) * S.h[a];
    u[a] = S.ud[0] ? S.ud[a][p] : static_cast<double>(S.uf[a][p]);
  }
  q = S.qd ? S.qd[p] : static_cast<double>(S.qf[p]);
  m = S.md ? S.md[p] : static_cast<double>(S.mf[p]);
  return true;
}

// meanVelocity (eb_initial_field.cpp:629-668), distributionResolution's
// moments (704-720), the total charge and the total |charge|: 12 sums.
__global__ void __launch_bounds__(kThreads) kSourceSums(Source S, DD* partial) {
  DD v[12];
  for (int p = blockIdx.x * kThreads + threadIdx.x; p < S.n; p += gridDim.x * kThreads) {
    double x[3], u[3], q, m;
    if (!sourceAt(S, p, x, u, q, m)) continue;
    const double gamma = gammaFromU(u[0], u[1], u[2]);
    v[0] = ddAdd(v[0], m);
    for (int a = 0; a < 3; ++a) {
      v[1 + a] = ddAddProduct(v[1 + a], m, u[a] / gamma);
      v[4 + a] = ddAddProduct(v[4 + a], m, x[a]);
      // m x^2 = (p + e) x with p + e = m x exactly
      const double p = m * x[a];
      v[7 + a] = ddAddProduct(v[7 + a], p, x[a]);
      v[7 + a] = ddAdd(v[7 + a], fma(m, x[a], -p) * x[a]);
    }
    v[10] = ddAdd(v[10], q);
    v[11] = ddAdd(v[11], fabs(q));
  }
  blockDD<12>(v, partial);
}

struct Lattice {
  int nx, ny, nz;           // cells of the field grid
  int pad;                  // transverse Poisson padding (cells)
  int sx, sy;               // slab vertices per row / column: nx + 2 pad + 1, ny + 2 pad + 1
  int lo[3], hi[3];         // Poisson domain (vertex indices, exclusive)
  double o[3], h[3];
};
__device__ __forceinline__ size_t slabAt(const Lattice& L, int i, int j, int p) {
  return (static_cast<size_t>(p) * L.sy + j) * L.sx + i;
}
__device__ __forceinline__ bool interior(const Lattice& L, int i, int j, int k) {
  return i > L.lo[0] && i < L.hi[0] && j > L.lo[1] && j < L.hi[1] && k > L.lo[2] && k < L.hi[2];
}

// CIC charge (depositCharge, eb_deposition.cpp:168-193) in 64-bit fixed
// point: pass 1 |taps| on a coarse exponent, pass 2 the signed taps on the
// exponent of twice the largest vertex; out-of-domain positions flag *err.
__global__ void kCic(Source S, Lattice L, int shift, int absolute, unsigned long long* acc, unsigned* err) {
  const int cells[3] = {L.nx, L.ny, L.nz};
  const double inverseVolume = 1.0 / (L.h[0] * L.h[1] * L.h[2]);
  const int vx = L.nx + 1, vy = L.ny + 1;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < S.n; p += gridDim.x * blockDim.x) {
    double x[3], u[3], q, m;
    if (!sourceAt(S, p, x, u, q, m)) continue;
    int cell[3];
    double f[3];
    bool inside = true;
    for (int a = 0; a < 3; ++a) {
      const double extent = static_cast<double>(cells[a]) * L.h[a];
      const double relative = x[a] - L.o[a];
      const double tolerance = 64.0 * kEpsilon * fmax(1.0, fmax(fabs(L.o[a]), extent));
      if (!isfinite(x[a]) || relative < -tolerance || relative > extent + tolerance) inside = false;
      const double c = fmax(0.0, fmin(static_cast<double>(cells[a]), (x[a] - L.o[a]) / L.h[a]));
      int k;
      if (c <= 0.0) k = 0;
      else if (c >= static_cast<double>(cells[a])) k = cells[a] - 1;
      else k = static_cast<int>(floor(c));
      cell[a] = k;
      f[a] = fmax(0.0, fmin(1.0, c - static_cast<double>(k)));
    }
    if (!inside) {
      atomicOr(err, 1u);
      continue;
    }
    for (int c = 0; c < 2; ++c)
      for (int b = 0; b < 2; ++b)
        for (int a = 0; a < 2; ++a) {
          const double w = q * inverseVolume * (a ? f[0] : 1.0 - f[0]) * (b ? f[1] : 1.0 - f[1]) * (c ? f[2] : 1.0 - f[2]);
          const size_t idx = (static_cast<size_t>(cell[2] + c) * vy + (cell[1] + b)) * vx + (cell[0] + a);
          atomicAdd(acc + idx, static_cast<unsigned long long>(llrint(ldexp(absolute ? fabs(w) : w, shift))));
        }
  }
}

__global__ void __launch_bounds__(kThreads) kMaxClear(unsigned long long* acc, size_t n, unsigned long long* out) {
  __shared__ unsigned long long sm[kThreads];
  unsigned long long m = 0ull;
  for (size_t t = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x; t < n;
       t += static_cast<size_t>(gridDim.x) * kThreads) {
    m = max(m, acc[t]);
    acc[t] = 0ull;
  }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) sm[threadIdx.x] = max(sm[threadIdx.x], sm[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, sm[0]);
}

__global__ void kToDouble(unsigned long long* acc, double* out, size_t n, int shift) {
  for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<size_t>(gridDim.x) * blockDim.x) {
    out[t] = ldexp(static_cast<double>(static_cast<long long>(acc[t])), -shift);
    acc[t] = 0ull;
  }
}

// The source (rho / epsilon0 at interior vertices of the padded lattice, 0
// elsewhere) into residual and direction, potential and image zero; the
// interior charge sum (eb_initial_field.cpp:981-1000).
__global__ void __launch_bounds__(kThreads) kSource(Lattice L, const double* rho, double* potential, double* residual,
                                                   double* direction, double* image, double vertexVolume, DD* partial) {
  const size_t total = static_cast<size_t>(L.sx) * L.sy * (L.nz + 2);
  DD charge;
  for (size_t t = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x; t < total;
       t += static_cast<size_t>(gridDim.x) * kThreads) {
    const int i = static_cast<int>(t % L.sx), j = static_cast<int>((t / L.sx) % L.sy), p = static_cast<int>(t / (static_cast<size_t>(L.sx) * L.sy));
    double r = 0.0;
    const int k = p - 1;
    if (p >= 1 && p <= L.nz && interior(L, i, j, k)) {
      const int ui = i - L.pad, uj = j - L.pad;
      const double v = (ui >= 0 && ui <= L.nx && uj >= 0 && uj <= L.ny)
                           ? rho[(static_cast<size_t>(k) * (L.ny + 1) + uj) * (L.nx + 1) + ui]
                           : 0.0;
      r = v / kEps0;
      charge = ddAddProduct(charge, v, vertexVolume);
    }
    potential[t] = 0.0;
    residual[t] = r;
    direction[t] = r;
    image[t] = 0.0;
  }
  DD v[1] = {charge};
  blockDD<1>(v, partial);
}

struct Stencil {
  double s[27];  // [z][y][x]
  double cx, cy, cz, diagonal;
  int cowan;
};

// image = A direction on interior vertices (0 elsewhere) and the partial
// sums of direction . image (applyNegativeLaplacian /
// applyNegativeCowanLaplacian, eb_initial_field.cpp:189-215, 281-316).
__global__ void __launch_bounds__(kThreads) kApply(Lattice L, Stencil A, const double* d, double* image, DD* partial) {
  const size_t total = static_cast<size_t>(L.sx) * L.sy * (L.nz + 2);
  DD acc;
  for (size_t t = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x; t < total;
       t += static_cast<size_t>(gridDim.x) * kThreads) {
    const int i = static_cast<int>(t % L.sx), j = static_cast<int>((t / L.sx) % L.sy), p = static_cast<int>(t / (static_cast<size_t>(L.sx) * L.sy));
    double out = 0.0;
    if (p >= 1 && p <= L.nz && interior(L, i, j, p - 1)) {
      if (A.cowan) {
        double value = 0.0;
        for (int z = 0; z < 3; ++z)
          for (int y = 0; y < 3; ++y)
            for (int x = 0; x < 3; ++x) value += A.s[(z * 3 + y) * 3 + x] * d[slabAt(L, i - 1 + x, j - 1 + y, p - 1 + z)];
        out = value;
      } else {
        out = A.diagonal * d[t] - A.cx * (d[slabAt(L, i - 1, j, p)] + d[slabAt(L, i + 1, j, p)]) -
              A.cy * (d[slabAt(L, i, j - 1, p)] + d[slabAt(L, i, j + 1, p)]) -
              A.cz * (d[slabAt(L, i, j, p - 1)] + d[slabAt(L, i, j, p + 1)]);
      }
      acc = ddAddProduct(acc, d[t], out);
    }
    image[t] = out;
  }
  DD v[1] = {acc};
  blockDD<1>(v, partial);
}

// potential += alpha direction, residual -= alpha image on interior
// vertices (updateInterior, eb_initial_field.cpp:409-424) and the partial
// sums of residual . residual.
__global__ void __launch_bounds__(kThreads) kUpdate(Lattice L, double alpha, double* potential, double* residual,
                                                   const double* d, const double* image, DD* partial) {
  const size_t total = static_cast<size_t>(L.sx) * L.sy * (L.nz + 2);
  DD acc;
  for (size_t t = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x; t < total;
       t += static_cast<size_t>(gridDim.x) * kThreads) {
    const int i = static_cast<int>(t % L.sx), j = static_cast<int>((t / L.sx) % L.sy), p = static_cast<int>(t / (static_cast<size_t>(L.sx) * L.sy));
    if (p >= 1 && p <= L.nz && interior(L, i, j, p - 1)) {
      potential[t] += alpha * d[t];
      const double r = residual[t] + (-alpha) * image[t];
      residual[t] = r;
      acc = ddAddProduct(acc, r, r);
    }
  }
  DD v[1] = {acc};
  blockDD<1>(v, partial);
}

// combineDirection (eb_initial_field.cpp:426-440) over the owned planes.
__global__ void kCombine(Lattice L, double beta, const double* residual, double* d) {
  const size_t total = static_cast<size_t>(L.sx) * L.sy * L.nz;
  for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < total;
       t += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t s = t + static_cast<size_t>(L.sx) * L.sy;  // planes 1..nz
    const int i = static_cast<int>(t % L.sx), j = static_cast<int>((t / L.sx) % L.sy), k = static_cast<int>(t / (static_cast<size_t>(L.sx) * L.sy));
    d[s] = interior(L, i, j, k) ? residual[s] + beta * d[s] : 0.0;
  }
}

struct CowanSmooth {
  double alpha[3], beta[3], delta[3];  // delta per axis: dYZ, dZX, dXY
  double translation;
};

// cowanSmoothedPotential minus the translation term on axis z, interior
// vertices only (eb_initial_field.cpp:318-352, 527-540).
__global__ void kSmooth(Lattice L, CowanSmooth c, int axis, const double* phi, double* out) {
  const size_t total = static_cast<size_t>(L.sx) * L.sy * (L.nz + 2);
  for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < total;
       t += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int i = static_cast<int>(t % L.sx), j = static_cast<int>((t / L.sx) % L.sy), p = static_cast<int>(t / (static_cast<size_t>(L.sx) * L.sy));
    double v = 0.0;
    if (p >= 1 && p <= L.nz && interior(L, i, j, p - 1)) {
      const int ta = (axis + 1) % 3, tb = (axis + 2) % 3;
      auto at = [&](const int q[3]) { return phi[slabAt(L, q[0], q[1], q[2])]; };
      const int c0[3] = {i, j, p};
      double value = c.alpha[axis] * at(c0);
      for (int sign = -1; sign <= 1; sign += 2) {
        int a[3] = {i, j, p}, b[3] = {i, j, p};
        a[ta] += sign;
        b[tb] += sign;
        value += c.beta[ta] * at(a);
        value += c.beta[tb] * at(b);
      }
      for (int sa = -1; sa <= 1; sa += 2)
        for (int sb = -1; sb <= 1; sb += 2) {
          int q[3] = {i, j, p};
          q[ta] += sa;
          q[tb] += sb;
          value += c.delta[axis] * at(q);
        }
      v = value - (axis == 2 ? c.translation * phi[t] : 0.0);
    }
    out[t] = v;
  }
}

// E component `axis` = -(s(+1) - s(0)) / h_axis (times lc for the plain
// Ez) from the slab s (eb_initial_field.cpp:449-467, 541-550).
__global__ void kGradient(Lattice L, int axis, double scale, const double* s, double* e, int n0, int n1, int n2) {
  const size_t total = static_cast<size_t>(n0) * n1 * n2;
  for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < total;
       t += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int i = static_cast<int>(t % n0), j = static_cast<int>((t / n0) % n1), k = static_cast<int>(t / (static_cast<size_t>(n0) * n1));
    const double hi = s[slabAt(L, i + L.pad + (axis == 0), j + L.pad + (axis == 1), k + 1 + (axis == 2))];
    const double lo = s[slabAt(L, i + L.pad, j + L.pad, k + 1)];
    e[t] = scale == 1.0 ? -(hi - lo) / L.h[axis] : -scale * (hi - lo) / L.h[axis];
  }
}

// B at -dt/2 by rigid translation (eb_initial_field.cpp:475-493, 553-574).
__global__ void kMagnetic(int nx, int ny, int nz, double scale, double lower, double upper, const double* ex,
                          const double* ey, double* bx, double* by, double* bz) {
  const size_t nBx = static_cast<size_t>(nx + 1) * ny * nz, nBy = static_cast<size_t>(nx) * (ny + 1) * nz,
               nBz = static_cast<size_t>(nx) * ny * (nz + 1);
  for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < nBx + nBy + nBz;
       t += static_cast<size_t>(gridDim.x) * blockDim.x) {
    if (t < nBx) {  // Bx (nx+1, ny, nz) from Ey (nx+1, ny, nz+1)
      const int i = static_cast<int>(t % (nx + 1)), j = static_cast<int>((t / (nx + 1)) % ny), k = static_cast<int>(t / (static_cast<size_t>(nx + 1) * ny));
      const size_t a = (static_cast<size_t>(k) * ny + j) * (nx + 1) + i, b = a + static_cast<size_t>(ny) * (nx + 1);
      bx[t] = -scale * (lower * ey[a] + upper * ey[b]);
    } else if (t < nBx + nBy) {  // By (nx, ny+1, nz) from Ex (nx, ny+1, nz+1)
      const size_t u = t - nBx;
      const int i = static_cast<int>(u % nx), j = static_cast<int>((u / nx) % (ny + 1)), k = static_cast<int>(u / (static_cast<size_t>(nx) * (ny + 1)));
      const size_t a = (static_cast<size_t>(k) * (ny + 1) + j) * nx + i, b = a + static_cast<size_t>(ny + 1) * nx;
      by[u] = scale * (lower * ex[a] + upper * ex[b]);
    } else {
      bz[t - nBx - nBy] = 0.0;
    }
  }
}

// clearComponentOutsidePhysical (eb_initial_field.cpp:578-604).
__global__ void kClearOutside(double* c, int n0, int n1, int n2, int electric, int axisC, int gx, int gy, int gz, int nx,
                              int ny, int nz) {
  const size_t total = static_cast<size_t>(n0) * n1 * n2;
  const int guard[3] = {gx, gy, gz}, cells[3] = {nx, ny, nz};
  for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < total;
       t += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int co[3] = {static_cast<int>(t % n0), static_cast<int>((t / n0) % n1), static_cast<int>(t / (static_cast<size_t>(n0) * n1))};
    bool inside = true;
    for (int axis = 0; axis < 3; ++axis) {
      const bool vertex = electric ? axis != axisC : axis == axisC;
      inside = inside && co[axis] >= guard[axis] &&
               (vertex ? co[axis] <= cells[axis] - guard[axis] : co[axis] < cells[axis] - guard[axis]);
    }
    if (!inside) c[t] = 0.0;
  }
}

struct Six {
  const double* c[6];
};

// validateGauss (eb_initial_field.cpp:749-811) on one rank: sums of error^2
// and source^2, max |error|, over the physical interior vertices.
__global__ void __launch_bounds__(kThreads) kGaussCheck(Six F, const double* rho, int nx, int ny, int nz, int lo0,
                                                       int lo1, int lo2, int hi0, int hi1, int hi2, double dx, double dy,
                                                       double dz, DD* partial, unsigned long long* maximum) {
  DD e2, s2;
  const size_t total = static_cast<size_t>(nx + 1) * (ny + 1) * nz;
  for (size_t t = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x; t < total;
       t += static_cast<size_t>(gridDim.x) * kThreads) {
    const int i = static_cast<int>(t % (nx + 1)), j = static_cast<int>((t / (nx + 1)) % (ny + 1)), k = static_cast<int>(t / (static_cast<size_t>(nx + 1) * (ny + 1)));
    if (!(k > lo2 && k < hi2 && j > lo1 && j < hi1 && i > lo0 && i < hi0)) continue;
    auto ex = [&](int a, int b, int c) { return F.c[0][(static_cast<size_t>(c) * (ny + 1) + b) * nx + a]; };
    auto ey = [&](int a, int b, int c) { return F.c[1][(static_cast<size_t>(c) * ny + b) * (nx + 1) + a]; };
    auto ez = [&](int a, int b, int c) { return F.c[2][(static_cast<size_t>(c) * (ny + 1) + b) * (nx + 1) + a]; };
    const double leftZ = k > 0 ? ez(i, j, k - 1) : 0.0;
    const double divergence = (ex(i, j, k) - ex(i - 1, j, k)) / dx + (ey(i, j, k) - ey(i, j - 1, k)) / dy +
                              (ez(i, j, k) - leftZ) / dz;
    const double source = rho[(static_cast<size_t>(k) * (ny + 1) + j) * (nx + 1) + i] / kEps0;
    const double error = divergence - source;
    e2 = ddAddProduct(e2, error, error);
    s2 = ddAddProduct(s2, source, source);
    atomicMaxAbs(maximum, error);
  }
  DD v[2] = {e2, s2};
  blockDD<2>(v, partial);
}

// initialFieldEnergy (eb_initial_field.cpp:670-696) with radiationSampleCell
// (eb_field.cpp:693-713): sums of |E|^2 and |B|^2 over the cells.
__global__ void __launch_bounds__(kThreads) kEnergies(Six F, int nx, int ny, int nz, DD* partial) {
  DD se, sb;
  const size_t total = static_cast<size_t>(nx) * ny * nz;
  for (size_t t = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x; t < total;
       t += static_cast<size_t>(gridDim.x) * kThreads) {
    const int i = static_cast<int>(t % nx), j = static_cast<int>((t / nx) % ny), k = static_cast<int>(t / (static_cast<size_t>(nx) * ny));
    auto at = [&](int c, int n0, int n1, int a, int b, int z) { return F.c[c][(static_cast<size_t>(z) * n1 + b) * n0 + a]; };
    double e[3], b[3];
    e[0] = 0.25 * (at(0, nx, ny + 1, i, j, k) + at(0, nx, ny + 1, i, j + 1, k) + at(0, nx, ny + 1, i, j, k + 1) +
                   at(0, nx, ny + 1, i, j + 1, k + 1));
    e[1] = 0.25 * (at(1, nx + 1, ny, i, j, k) + at(1, nx + 1, ny, i + 1, j, k) + at(1, nx + 1, ny, i, j, k + 1) +
                   at(1, nx + 1, ny, i + 1, j, k + 1));
    e[2] = 0.25 * (at(2, nx + 1, ny + 1, i, j, k) + at(2, nx + 1, ny + 1, i + 1, j, k) +
                   at(2, nx + 1, ny + 1, i, j + 1, k) + at(2, nx + 1, ny + 1, i + 1, j + 1, k));
    b[0] = 0.5 * (at(3, nx + 1, ny, i, j, k) + at(3, nx + 1, ny, i + 1, j, k));
    b[1] = 0.5 * (at(4, nx, ny + 1, i, j, k) + at(4, nx, ny + 1, i, j + 1, k));
    b[2] = 0.5 * (at(5, nx, ny, i, j, k) + at(5, nx, ny, i, j, k + 1));
    se = ddAdd(se, e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    sb = ddAdd(sb, b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  }
  DD v[2] = {se, sb};
  blockDD<2>(v, partial);
}

template <class R>
__global__ void kStore(const double* src, R* dst, size_t n) {
  for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<size_t>(gridDim.x) * blockDim.x)
    dst[t] = static_cast<R>(src[t]);
}

long double toLong(const DD& v) { return static_cast<long double>(v.hi) + static_cast<long double>(v.lo); }

// The device scratch of one solve.
struct Work {
  DeviceArray<DD> partial, sums;
  DeviceArray<unsigned long long> acc, peak;
  DeviceArray<double> rho, slab[4], field[6];
  DeviceArray<unsigned> err;
};

// Sum of K partial accumulators of `blocks` blocks, read back.
template <int K>
void finish(hipStream_t st, Work& W, int blocks, DD out[K]) {
  kFinishDD<<<1, 64, 0, st>>>(W.partial.get(), blocks, K, W.sums.get());
  launchCheck("initial field sums");
  hipCheck(hipMemcpyAsync(out, W.sums.get(), K * sizeof(DD), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
  hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
}

unsigned gridFor(std::size_t n) { return static_cast<unsigned>(std::min<std::size_t>(std::max<std::size_t>((n + 255) / 256, 1), 4096)); }

// CIC charge of the sources on the (nx+1)(ny+1)(nz+1) vertices into W.rho.
void depositCic(hipStream_t st, Work& W, const Source& S, const Lattice& L, double totalAbsCharge) {
  const std::size_t nv = static_cast<std::size_t>(L.nx + 1) * (L.ny + 1) * (L.nz + 1);
  W.acc.resize(nv);
  W.rho.resize(nv);
  hipCheck(hipMemsetAsync(W.acc.get(), 0, nv * sizeof(unsigned long long), st), "hipMemsetAsync");
  hipCheck(hipMemsetAsync(W.err.get(), 0, sizeof(unsigned), st), "hipMemsetAsync");
  int shift = 0;
  if (S.n > 0 && totalAbsCharge > 0.0) {
    int e2 = 0;
    (void)std::frexp(totalAbsCharge / (L.h[0] * L.h[1] * L.h[2]), &e2);
    const int coarse = 62 - e2;
    hipCheck(hipMemsetAsync(W.peak.get(), 0, sizeof(unsigned long long), st), "hipMemsetAsync");
    kCic<<<gridFor(S.n), 256, 0, st>>>(S, L, coarse, 1, W.acc.get(), W.err.get());
    kMaxClear<<<kBlocks, kThreads, 0, st>>>(W.acc.get(), nv, W.peak.get());
    launchCheck("initial charge bound");
    unsigned long long peak = 0;
    unsigned err = 0;
    hipCheck(hipMemcpyAsync(&peak, W.peak.get(), sizeof(peak), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    hipCheck(hipMemcpyAsync(&err, W.err.get(), sizeof(err), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (err) throw std::out_of_range("Particle trajectory endpoint is outside the E/B domain");
    const double bound = std::ldexp(static_cast<double>(peak) + 1.0, -coarse);
    (void)std::frexp(2.0 * bound, &e2);
    shift = 62 - e2;
    kCic<<<gridFor(S.n), 256, 0, st>>>(S, L, shift, 0, W.acc.get(), W.err.get());
  }
  kToDouble<<<gridFor(nv), 256, 0, st>>>(W.acc.get(), W.rho.get(), nv, shift);
  launchCheck("initial charge");
}

template <class R>
void initializeGauss(FieldStateT<R>& F, const Source& S, const double origin[3], int model, double maxBetaT,
                     const std::size_t guard[3], double tolerance, std::size_t maxIterations, bool includeCpml,
                     aprl_gauss_report& report) {
  if (!(tolerance > 0.0) || !(tolerance < 1.0) || !std::isfinite(tolerance) || maxIterations == 0)
    throw std::invalid_argument("Invalid initial Gauss-field convergence controls");
  if (!(maxBetaT >= 0.0) || !(maxBetaT < 1.0) || !std::isfinite(maxBetaT))
    throw std::invalid_argument("Invalid initial self-field maximum mean transverse beta");
  if (model < 0 || model > 2) throw std::invalid_argument("Unknown initial self-field model");
  const bool cowanModel = model == 2;
  const aprl_grid_geometry& g = F.g;
  if (cowanModel && g.solver != 1)
    throw std::invalid_argument("cowan-relativistic-poisson requires the cowan-z Maxwell solver");
  if (includeCpml && (guard[0] < 2 || guard[1] < 2))
    throw std::invalid_argument(
        "Transverse CPML-source initialization requires at least two physical guard cells in x and y");
  const int pad = includeCpml ? 2 : 0;
  const std::size_t cells[3] = {g.nx, g.ny, g.nz};
  Lattice L{};
  L.nx = static_cast<int>(g.nx);
  L.ny = static_cast<int>(g.ny);
  L.nz = static_cast<int>(g.nz);
  L.pad = pad;
  L.sx = L.nx + 2 * pad + 1;
  L.sy = L.ny + 2 * pad + 1;
  L.h[0] = g.dx;
  L.h[1] = g.dy;
  L.h[2] = g.dz;
  for (int a = 0; a < 3; ++a) L.o[a] = origin[a];
  int physLo[3], physHi[3];
  for (int a = 0; a < 3; ++a) {
    if (2 * guard[a] + 2 >= cells[a])
      throw std::invalid_argument("Initial static-field guard leaves fewer than two physical cells");
    physLo[a] = static_cast<int>(guard[a]);
    physHi[a] = static_cast<int>(cells[a] - guard[a]);
    L.lo[a] = includeCpml && a < 2 ? 0 : static_cast<int>(guard[a]);
    const int pc = a == 0 ? L.nx + 2 * pad : (a == 1 ? L.ny + 2 * pad : L.nz);
    L.hi[a] = pc - L.lo[a];
  }
  hipStream_t st = F.stream;
  Work W;
  W.partial.resize(12 * kBlocks);
  W.sums.resize(12);
  W.peak.resize(1);
  W.err.resize(1);

  // mean velocity, moments and total charge
  kSourceSums<<<kBlocks, kThreads, 0, st>>>(S, W.partial.get());
  launchCheck("initial field source sums");
  DD sums[12];
  finish<12>(st, W, kBlocks, sums);
  const long double mass = toLong(sums[0]);
  if (!(mass > 0.0L)) throw std::runtime_error("Initial relativistic-Poisson field requires positive represented mass");
  const double betaX = static_cast<double>(toLong(sums[1]) / mass), betaY = static_cast<double>(toLong(sums[2]) / mass);
  const double betaZ = static_cast<double>(toLong(sums[3]) / mass);
  const double betaT = std::hypot(betaX, betaY);
  if (!std::isfinite(betaZ) || !std::isfinite(betaT) || betaT >= 1.0 || std::abs(betaZ) >= 1.0)
    throw std::runtime_error("Initial particle mean velocity is not a finite subluminal vector");
  if ((model == 0 || cowanModel) && betaT > maxBetaT) {
    std::ostringstream m;
    m.precision(12);
    m << (cowanModel ? "cowan-relativistic-poisson" : "relativistic-poisson") << " mean beta_box_transverse=" << betaT
      << " exceeds initial_self_field.maximum_mean_transverse_beta=" << maxBetaT
      << "; the rigid-beam initial field assumes a mean velocity along the boost z axis. Align/recenter the beam or "
         "explicitly relax this approximation limit.";
    throw std::runtime_error(m.str());
  }
  const double applied = (model == 0 || cowanModel) ? betaZ : 0.0;
  const double lc = std::fma(-applied, applied, 1.0);
  if (!cowanModel && !(lc > 64.0 * kEpsilon))
    throw std::runtime_error(
        "Initial relativistic-Poisson operator is ill-conditioned because the bunch remains too relativistic in the "
        "simulation frame; choose boost_gamma closer to the bunch mean gamma");
  // operators (CowanPoissonOperator, eb_initial_field.cpp:209-278)
  Stencil A{};
  CowanSmooth smooth{};
  const double spacing[3] = {g.dx, g.dy, g.dz};
  if (cowanModel) {
    A.cowan = 1;
    const CowanCoef& c = F.cw;
    const double velocity = applied * kC;
    const double translation = velocity * velocity * kEps0 * kMu0;
    const double deltas[3] = {c.deltaYZ, c.deltaZX, c.deltaXY};
    double st3[3][3][3] = {};
    for (int axis = 0; axis < 3; ++axis) {
      const int ta = (axis + 1) % 3, tb = (axis + 2) % 3;
      const double delta = deltas[axis];
      double minimum = std::numeric_limits<double>::infinity();
      for (int a = -1; a <= 1; a += 2)
        for (int b = -1; b <= 1; b += 2)
          minimum = std::min(minimum, c.alpha[axis] + 2.0 * c.beta[ta] * a + 2.0 * c.beta[tb] * b + 4.0 * delta * a * b);
      if (axis == 2) minimum -= translation;
      if (!(minimum > 64.0 * kEpsilon)) {
        std::ostringstream m;
        m.precision(12);
        m << "Initial Cowan-Poisson operator fails the sufficient positive-definiteness bound on axis " << axis
          << ": minimum smoothing symbol" << (axis == 2 ? " minus v^2*epsilon0*mu0" : "") << "=" << minimum
          << "; choose boost_gamma closer to the bunch mean gamma or a grid aspect ratio with positive Cowan symbols";
        throw std::runtime_error(m.str());
      }
      const double inv2 = 1.0 / (spacing[axis] * spacing[axis]);
      for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b) {
          const double weight = a == 0 ? (b == 0 ? c.alpha[axis] : c.beta[tb]) : (b == 0 ? c.beta[ta] : delta);
          int point[3] = {1, 1, 1};
          point[ta] += a;
          point[tb] += b;
          st3[point[2]][point[1]][point[0]] += 2.0 * weight * inv2;
          --point[axis];
          st3[point[2]][point[1]][point[0]] -= weight * inv2;
          point[axis] += 2;
          st3[point[2]][point[1]][point[0]] -= weight * inv2;
        }
      smooth.alpha[axis] = c.alpha[axis];
      smooth.beta[axis] = c.beta[axis];
      smooth.delta[axis] = deltas[axis];
    }
    const double tz = translation / (g.dz * g.dz);
    st3[1][1][1] -= 2.0 * tz;
    st3[0][1][1] += tz;
    st3[2][1][1] += tz;
    std::memcpy(A.s, st3, sizeof(A.s));
    smooth.translation = translation;
  } else {
    A.cx = 1.0 / (g.dx * g.dx);
    A.cy = 1.0 / (g.dy * g.dy);
    A.cz = lc / (g.dz * g.dz);
    A.diagonal = 2.0 * (A.cx + A.cy + A.cz);
  }
  report.mean_beta_z = betaZ;
  report.mean_beta_transverse = betaT;
  report.transverse_poisson_padding_cells = static_cast<uint64_t>(pad);
  for (int a = 0; a < 3; ++a) report.static_field_guard_cells[a] = static_cast<uint64_t>(L.lo[a]);
  // distributionResolution (698-747)
  {
    const double porigin[3] = {origin[0] - static_cast<double>(pad) * g.dx, origin[1] - static_cast<double>(pad) * g.dy,
                               origin[2]};
    for (int a = 0; a < 3; ++a) {
      const long double mean = toLong(sums[4 + a]) / mass;
      const long double variance = std::max(0.0L, toLong(sums[7 + a]) / mass - mean * mean);
      const double rms = static_cast<double>(std::sqrt(variance));
      report.rms_position_cells[a] = rms / spacing[a];
      if (rms > 0.0) {
        const double lower = porigin[a] + static_cast<double>(L.lo[a]) * spacing[a];
        const double upper = porigin[a] + static_cast<double>(L.hi[a]) * spacing[a];
        report.centre_padding_rms[a] = std::min(static_cast<double>(mean) - lower, upper - static_cast<double>(mean)) / rms;
      } else {
        report.centre_padding_rms[a] = 0.0;
      }
    }
  }
  const long double totalCharge = toLong(sums[10]);
  report.total_charge = static_cast<double>(totalCharge);

  // charge, source and the interior-charge check (981-1007); the total
  // |charge| bounds every vertex for the coarse fixed-point pass
  depositCic(st, W, S, L, static_cast<double>(toLong(sums[11])));
  const std::size_t slabN = static_cast<std::size_t>(L.sx) * L.sy * (L.nz + 2);
  for (auto& s : W.slab) s.resize(slabN);
  double* potential = W.slab[0].get();
  double* residual = W.slab[1].get();
  double* direction = W.slab[2].get();
  double* image = W.slab[3].get();
  const double vertexVolume = g.dx * g.dy * g.dz;
  kSource<<<kBlocks, kThreads, 0, st>>>(L, W.rho.get(), potential, residual, direction, image, vertexVolume,
                                        W.partial.get());
  launchCheck("initial field source");
  DD interiorCharge[1];
  finish<1>(st, W, kBlocks, interiorCharge);
  if (std::abs((toLong(interiorCharge[0]) - totalCharge) / totalCharge) > 1.0e-12L)
    throw std::runtime_error(
        "Initial CIC charge cloud touches the static zero-potential boundary; move every particle at least one "
        "complete cell inside the CPML entrance or enlarge the physical mesh padding");

  // conjugate gradients (1010-1063)
  kUpdate<<<kBlocks, kThreads, 0, st>>>(L, 0.0, potential, residual, direction, image, W.partial.get());
  launchCheck("initial field source norm");
  DD norm[1];
  finish<1>(st, W, kBlocks, norm);
  const long double source2 = toLong(norm[0]);
  if (!(source2 > 0.0L) || !std::isfinite(static_cast<double>(source2)))
    throw std::runtime_error("Initial particle charge produced an invalid Poisson source norm");
  long double r2 = source2;
  double relative = 0.0;
  std::size_t iterations = 0;
  for (std::size_t it = 0; it < maxIterations; ++it) {
    kApply<<<kBlocks, kThreads, 0, st>>>(L, A, direction, image, W.partial.get());
    launchCheck("initial field operator");
    DD den[1];
    finish<1>(st, W, kBlocks, den);
    const long double d = toLong(den[0]);
    if (!(d > 0.0L) || !std::isfinite(static_cast<double>(d)))
      throw std::runtime_error("Initial Gauss-field CG lost positive definiteness");
    const double alpha = static_cast<double>(r2 / d);
    kUpdate<<<kBlocks, kThreads, 0, st>>>(L, alpha, potential, residual, direction, image, W.partial.get());
    launchCheck("initial field update");
    DD nxt[1];
    finish<1>(st, W, kBlocks, nxt);
    const long double next = toLong(nxt[0]);
    iterations = it + 1;
    relative = static_cast<double>(std::sqrt(next / source2));
    if (relative <= tolerance) {
      r2 = next;
      break;
    }
    if (!(next >= 0.0L) || !std::isfinite(static_cast<double>(next)))
      throw std::runtime_error("Initial Gauss-field CG produced a non-finite residual");
    const double beta = static_cast<double>(next / r2);
    kCombine<<<gridFor(static_cast<std::size_t>(L.sx) * L.sy * L.nz), 256, 0, st>>>(L, beta, residual, direction);
    launchCheck("initial field direction");
    r2 = next;
  }
  report.iterations = iterations;
  report.relative_residual = relative;
  if (relative > tolerance)
    throw std::runtime_error(
        "Initial Gauss-field solve did not converge; increase initial_self_field.maximum_iterations, relax its "
        "relative_tolerance, or reduce the grid aspect ratio");

  // fields from the potential, in FP64 scratch of the component shapes
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const int shapes[6][3] = {{nx, ny + 1, nz + 1}, {nx + 1, ny, nz + 1}, {nx + 1, ny + 1, nz},
                            {nx + 1, ny, nz},     {nx, ny + 1, nz},     {nx, ny, nz + 1}};
  for (int a = 0; a < 6; ++a) W.field[a].resize(static_cast<std::size_t>(shapes[a][0]) * shapes[a][1] * shapes[a][2]);
  for (int axis = 0; axis < 3; ++axis) {
    const std::size_t n = W.field[axis].size();
    if (cowanModel) {
      kSmooth<<<gridFor(slabN), 256, 0, st>>>(L, smooth, axis, potential, image);
      kGradient<<<gridFor(n), 256, 0, st>>>(L, axis, 1.0, image, W.field[axis].get(), shapes[axis][0], shapes[axis][1],
                                            shapes[axis][2]);
    } else {
      kGradient<<<gridFor(n), 256, 0, st>>>(L, axis, axis == 2 ? lc : 1.0, potential, W.field[axis].get(),
                                            shapes[axis][0], shapes[axis][1], shapes[axis][2]);
    }
    launchCheck("initial electric field");
  }
  const double magneticScale = cowanModel ? applied * kC * kEps0 * kMu0 : applied / kC;
  const double upper = 0.5 + 0.5 * applied * kC * g.dt / g.dz;
  const double lower = 1.0 - upper;
  if (upper < 0.0 || upper > 1.0) throw std::runtime_error("Initial B half-step shift exceeds one longitudinal cell");
  const std::size_t nB = W.field[3].size() + W.field[4].size() + W.field[5].size();
  kMagnetic<<<gridFor(nB), 256, 0, st>>>(nx, ny, nz, magneticScale, lower, upper, W.field[0].get(), W.field[1].get(),
                                         W.field[3].get(), W.field[4].get(), W.field[5].get());
  launchCheck("initial magnetic field");
  if (includeCpml)
    for (int a = 0; a < 6; ++a) {
      kClearOutside<<<gridFor(W.field[a].size()), 256, 0, st>>>(
          W.field[a].get(), shapes[a][0], shapes[a][1], shapes[a][2], a < 3 ? 1 : 0, a % 3, static_cast<int>(guard[0]),
          static_cast<int>(guard[1]), static_cast<int>(guard[2]), nx, ny, nz);
      launchCheck("initial field CPML clear");
    }
  // validateGauss with the unpadded charge (1082-1089)
  Six six{};
  for (int a = 0; a < 6; ++a) six.c[a] = W.field[a].get();
  hipCheck(hipMemsetAsync(W.peak.get(), 0, sizeof(unsigned long long), st), "hipMemsetAsync");
  kGaussCheck<<<kBlocks, kThreads, 0, st>>>(six, W.rho.get(), nx, ny, nz, physLo[0], physLo[1], physLo[2], physHi[0],
                                            physHi[1], physHi[2], g.dx, g.dy, g.dz, W.partial.get(), W.peak.get());
  launchCheck("initial Gauss check");
  DD gauss[2];
  finish<2>(st, W, kBlocks, gauss);
  unsigned long long maxBits = 0;
  hipCheck(hipMemcpy(&maxBits, W.peak.get(), sizeof(maxBits), hipMemcpyDeviceToHost), "hipMemcpy");
  double maximum;
  std::memcpy(&maximum, &maxBits, sizeof(maximum));
  report.maximum_gauss_residual = maximum;
  const long double s2 = toLong(gauss[1]);
  report.relative_residual = s2 > 0.0L ? static_cast<double>(std::sqrt(toLong(gauss[0]) / s2)) : 0.0;
  if (report.relative_residual > 4.0 * tolerance)
    throw std::runtime_error("Initial E field failed the post-solve discrete Gauss check");
  kEnergies<<<kBlocks, kThreads, 0, st>>>(six, nx, ny, nz, W.partial.get());
  launchCheck("initial field energies");
  DD energy[2];
  finish<2>(st, W, kBlocks, energy);
  const long double volume = static_cast<long double>(g.dx) * static_cast<long double>(g.dy) * static_cast<long double>(g.dz);
  report.initial_electric_energy = static_cast<double>(0.5L * static_cast<long double>(kEps0) * toLong(energy[0]) * volume);
  report.initial_magnetic_energy = static_cast<double>(0.5L / static_cast<long double>(kMu0) * toLong(energy[1]) * volume);
  // store E and B (clearFields: the grid's E and B are these fields)
  for (int a = 0; a < 6; ++a) {
    if (F.compSize(a) != W.field[a].size()) throw std::logic_error("Initial field component shape mismatch");
    kStore<R><<<gridFor(W.field[a].size()), 256, 0, st>>>(W.field[a].get(), F.base(a), W.field[a].size());
  }
  launchCheck("initial field store");
  hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
}

}  // namespace
}  // namespace aprl_gpu

using namespace aprl_gpu;

extern "C" {

aprl_status aprl_fields_initialize_gauss(aprl_fields* fields, aprl_particles* particles, const double* origin,
                                         aprl_initial_model model, double max_mean_transverse_beta,
                                         const size_t* guard_cells, double relative_tolerance,
                                         size_t maximum_iterations, int include_transverse_cpml_particles,
                                         aprl_gauss_report* report) {
  return guarded([&] {
    if (!fields || !particles || !origin || !guard_cells || !report)
      throw std::invalid_argument("Null field grid, particles, origin, guard or report");
    if (particles->fields != fields) throw std::invalid_argument("Particle store and field grid do not belong together");
    aprl_gauss_report r{};
    Source S{};
    if (fields->fp32) {
      SplitStore& P = particles->sp;
      const FieldState32& F = fields->s32;
      const double h[3] = {F.g.dx, F.g.dy, F.g.dz};
      for (int a = 0; a < 3; ++a) {
        S.cell[a] = P.cell[a].get();
        S.off[a] = P.s[a].get();
        S.h[a] = h[a];
        S.uf[a] = P.s[3 + a].get();
      }
      S.qf = P.s[6].get();
      S.mf = P.s[7].get();
      S.state = P.state.get();
      S.n = static_cast<int>(P.count);
      initializeGauss(fields->s32, S, origin, static_cast<int>(model), max_mean_transverse_beta, guard_cells,
                      relative_tolerance, maximum_iterations, include_transverse_cpml_particles != 0, r);
    } else {
      ParticleStore& P = particles->p;
      for (int a = 0; a < 3; ++a) {
        S.x[a] = P.s[a].get();
        S.ud[a] = P.s[3 + a].get();
      }
      S.qd = P.s[6].get();
      S.md = P.s[7].get();
      S.state = P.state.get();
      S.n = static_cast<int>(P.count);
      initializeGauss(fields->s, S, origin, static_cast<int>(model), max_mean_transverse_beta, guard_cells,
                      relative_tolerance, maximum_iterations, include_transverse_cpml_particles != 0, r);
    }
    *report = r;
  });
}

}  // extern "C"
