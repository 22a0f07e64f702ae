// Oracle: the moving window's inlet reference (TEST INFRASTRUCTURE ONLY;
// never linked into the product).
//
// Restates, for one rank and in memory (no block files):
//   blocked_cowan_reference.cpp:315-384  configure / validate: guards, interior nodes, native shapes
//   blocked_cowan_reference.cpp:505-551  makeModes: the 27-point Cowan stencil with the
//                                        translation term, diagonalized by the x/y sine modes
//   blocked_cowan_reference.cpp:559-655  the source charge: CIC (deposit) or the production
//                                        longitudinal basis with linear / x-binomial3 transverse
//                                        weights (depositQuintic), accumulated per node in the
//                                        reference's particle order
//   blocked_cowan_reference.cpp:709-770  the solve: rho / eps0, 2-D DST-I (FFTW RODFT00: y_k =
//                                        2 sum_j x_j sin(pi (j+1)(k+1) / (n+1)) per axis), one
//                                        Thomas sweep per mode along z, the inverse DST with
//                                        1 / (4 (nx+1) (ny+1))
//   blocked_cowan_reference.cpp:822-950  the native E(0), B(-dt/2) from the potential
//   simulation_moving_window.cpp:280-294 movingWindowFieldValue: the reference at
//                                        z = k - translationBeta step, linear in z
// The DST is evaluated by its definition (a direct sum), where the reference
// calls FFTW. The reference translation unit needs FFTW3, which is not in
// this image, so the restatement is pinned without it (tests/test_oracle_inlet_reference.py):
// the potential solves the 27-point operator to round-off, the direct DST
// is its own inverse up to 2(n+1) per axis, and on the same grid, source,
// beta and guards the fields equal those of the Cowan initializer
// (orc_initial_field.hpp, bitwise equal to the reference build) to its
// conjugate-gradient tolerance, as the reference documents
// (blocked_cowan_reference.h:68-70).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <vector>

#include "orc_particles.hpp"

namespace orc {
namespace inletref {

struct Config {
  std::size_t nx = 0, ny = 0, nz = 0;
  double dx = 0, dy = 0, dz = 0, dt = 0;
  double origin[3] = {0, 0, 0};
  double beta = 0;                 // the solver's mean held beta (mass weighted)
  std::size_t guards[3] = {0, 0, 0};
  std::size_t blockLayers = 64;    // only the shaped deposit's accumulation order depends on it
  int shape = -1;                  // -1 CIC, else kQuintic / kNotch / kBinomial3
  bool xBinomial = false;
};

struct Reference {
  Config c;
  Cowan coefficient;
  double translation = 0, upperWeight = 0, lowerWeight = 0, magneticScale = 0;
  std::size_t lower[3] = {0, 0, 0}, upper[3] = {0, 0, 0}, nodes[3] = {0, 0, 0};
  std::size_t area = 0, fullArea = 0, blockLayers = 0, blocks = 0;
  std::size_t nx[6] = {}, ny[6] = {}, nz[6] = {};
  std::vector<double> rho;       // interior nodes, x fastest (charge density, C/m^3)
  std::vector<double> phi;       // full vertex lattice (nx+1)(ny+1)(nz+1), zero outside the interior
  std::array<std::vector<double>, 6> field;  // native Ex Ey Ez Bx By Bz of the whole grid
  double minimumModeEigenvalue = 0, minimumPivot = 0;
  double sourceCharge = 0, depositedCharge = 0;

  bool interior(std::size_t i, std::size_t j, std::size_t k) const {
    return i > lower[0] && i < upper[0] && j > lower[1] && j < upper[1] && k > lower[2] && k < upper[2];
  }
  double& phiAt(std::size_t i, std::size_t j, std::size_t k) { return phi[k * fullArea + j * (c.nx + 1) + i]; }
  double phiAt(std::size_t i, std::size_t j, std::size_t k) const { return phi[k * fullArea + j * (c.nx + 1) + i]; }
  double value(unsigned component, std::size_t i, std::size_t j, std::size_t k) const {
    if (component >= 6 || i >= nx[component] || j >= ny[component] || k >= nz[component])
      throw std::out_of_range("Blocked field query outside native global Yee grid");
    return field[component][(k * ny[component] + j) * nx[component] + i];
  }
};

inline double deltaFor(const Cowan& c, unsigned axis) {
  return axis == 0 ? c.deltaYZ : (axis == 1 ? c.deltaZX : c.deltaXY);
}

// configure + validate, blocked_cowan_reference.cpp:315-384
inline void configure(Reference& r) {
  const Config& c = r.c;
  if (!std::isfinite(c.beta) || std::abs(c.beta) >= 1.0 || c.blockLayers == 0)
    throw std::invalid_argument("Invalid blocked Cowan reference configuration");
  if (c.xBinomial && c.shape != kBinomial3)
    throw std::invalid_argument("Transverse three-point reference requires longitudinal linear-binomial3");
  const std::size_t cells[3] = {c.nx, c.ny, c.nz};
  const double h[3] = {c.dx, c.dy, c.dz};
  for (unsigned a = 0; a < 3; ++a) {
    if (!std::isfinite(h[a]) || h[a] <= 0. || !std::isfinite(c.origin[a]) || cells[a] < 3 ||
        c.guards[a] > (cells[a] - 3) / 2)
      throw std::invalid_argument("Invalid blocked reference geometry or guard/padding");
    r.lower[a] = c.guards[a];
    r.upper[a] = cells[a] - c.guards[a];
    r.nodes[a] = r.upper[a] - r.lower[a] - 1;
  }
  if (!std::isfinite(c.dt) || c.dt <= 0. || !std::isfinite(c.dx * c.dy * c.dz) || c.dx * c.dy * c.dz <= 0.)
    throw std::invalid_argument("Invalid blocked reference time step or cell volume");
  r.blockLayers = std::min(c.blockLayers, r.nodes[2]);
  r.blocks = (r.nodes[2] - 1) / r.blockLayers + 1;
  r.area = r.nodes[0] * r.nodes[1];
  r.fullArea = (c.nx + 1) * (c.ny + 1);
  const std::size_t nxs[6] = {c.nx, c.nx + 1, c.nx + 1, c.nx + 1, c.nx, c.nx};
  const std::size_t nys[6] = {c.ny + 1, c.ny, c.ny + 1, c.ny, c.ny + 1, c.ny};
  for (unsigned a = 0; a < 6; ++a) {
    r.nx[a] = nxs[a];
    r.ny[a] = nys[a];
    r.nz[a] = c.nz + ((a == 0 || a == 1 || a == 5) ? 1 : 0);
  }
  r.coefficient = Cowan::forZDispersion(c.dx, c.dy, c.dz);
  const double speed = c.beta * kC;
  r.translation = speed * speed * kEps0 * kMu0;
  r.upperWeight = 0.5 + 0.5 * c.beta * kC * c.dt / c.dz;
  r.lowerWeight = 1. - r.upperWeight;
  r.magneticScale = c.beta * kC * kEps0 * kMu0;
  if (!(r.upperWeight >= 0. && r.upperWeight <= 1.))
    throw std::invalid_argument("Blocked initial B half-step shift exceeds one cell");
}

// makeModes, blocked_cowan_reference.cpp:505-551: the 27-point stencil and
// each (i, j) sine mode's constant tridiagonal along z.
inline void makeModes(Reference& r, double stencil[3][3][3], std::vector<double>& diagonal, std::vector<double>& off) {
  const Config& c = r.c;
  for (int z = 0; z < 3; ++z)
    for (int y = 0; y < 3; ++y)
      for (int x = 0; x < 3; ++x) stencil[z][y][x] = 0.;
  const double spacing[3] = {c.dx, c.dy, c.dz};
  for (unsigned axis = 0; axis < 3; ++axis) {
    const unsigned ta = (axis + 1) % 3, tb = (axis + 2) % 3;
    const double delta = deltaFor(r.coefficient, axis);
    const double inverse = 1. / (spacing[axis] * spacing[axis]);
    for (int a = -1; a <= 1; ++a)
      for (int b = -1; b <= 1; ++b) {
        const double weight = a == 0 ? (b == 0 ? r.coefficient.alpha[axis] : r.coefficient.beta[tb])
                                     : (b == 0 ? r.coefficient.beta[ta] : delta);
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
  const double translationZ = r.translation / (c.dz * c.dz);
  stencil[1][1][1] -= 2. * translationZ;
  stencil[0][1][1] += translationZ;
  stencil[2][1][1] += translationZ;
  diagonal.assign(r.area, 0.);
  off.assign(r.area, 0.);
  r.minimumModeEigenvalue = std::numeric_limits<double>::infinity();
  const double pi = std::acos(-1.);
  for (std::size_t j = 0; j < r.nodes[1]; ++j)
    for (std::size_t i = 0; i < r.nodes[0]; ++i) {
      const double tx = pi * (i + 1) / (r.nodes[0] + 1), ty = pi * (j + 1) / (r.nodes[1] + 1);
      double symbol[3] = {};
      for (unsigned z = 0; z < 3; ++z)
        for (unsigned y = 0; y < 3; ++y)
          for (unsigned x = 0; x < 3; ++x)
            symbol[z] += stencil[z][y][x] * std::cos((static_cast<int>(x) - 1) * tx) *
                         std::cos((static_cast<int>(y) - 1) * ty);
      const double scale = std::max(std::abs(symbol[1]), std::max(std::abs(symbol[0]), std::abs(symbol[2])));
      if (!std::isfinite(scale) || scale == 0. ||
          std::abs(symbol[0] - symbol[2]) > 128. * std::numeric_limits<double>::epsilon() * scale)
        throw std::runtime_error("Blocked Cowan modal matrix is not finite symmetric");
      const std::size_t m = j * r.nodes[0] + i;
      diagonal[m] = symbol[1];
      off[m] = 0.5 * (symbol[0] + symbol[2]);
      const double eigenvalue = diagonal[m] - 2. * std::abs(off[m]) * std::cos(pi / (r.nodes[2] + 1));
      if (!(eigenvalue > 0.) || !std::isfinite(eigenvalue))
        throw std::runtime_error("Blocked Cowan modal matrix is not positive definite");
      r.minimumModeEigenvalue = std::min(r.minimumModeEigenvalue, eigenvalue);
    }
}

// deposit (CIC), blocked_cowan_reference.cpp:595-655, per node in particle order.
inline void depositCic(Reference& r, std::size_t n, const double* x3, const double* q) {
  const Config& c = r.c;
  r.rho.assign(r.area * r.nodes[2], 0.);
  const double h[3] = {c.dx, c.dy, c.dz};
  const std::size_t cells[3] = {c.nx, c.ny, c.nz};
  const double volume = c.dx * c.dy * c.dz, inverseVolume = 1. / volume;
  long double source = 0.;
  for (std::size_t p = 0; p < n; ++p) {
    if (!std::isfinite(q[p])) throw std::invalid_argument("Nonfinite blocked source charge");
    source += q[p];
    std::size_t cell[3];
    double fraction[3];
    for (unsigned a = 0; a < 3; ++a) {
      const double relative = x3[3 * p + a] - c.origin[a], extent = cells[a] * h[a];
      const double tolerance =
          64. * std::numeric_limits<double>::epsilon() * std::max(1., std::max(std::abs(c.origin[a]), extent));
      if (!std::isfinite(x3[3 * p + a]) || relative < -tolerance || relative > extent + tolerance)
        throw std::out_of_range("Blocked source particle outside global E/B domain");
      const double normalized = std::max(0., std::min(static_cast<double>(cells[a]), relative / h[a]));
      cell[a] = normalized <= 0. ? 0
                                 : (normalized >= static_cast<double>(cells[a]) ? cells[a] - 1
                                                                                : static_cast<std::size_t>(std::floor(normalized)));
      fraction[a] = std::max(0., std::min(1., normalized - static_cast<double>(cell[a])));
    }
    for (unsigned cc = 0; cc < 2; ++cc)
      for (unsigned b = 0; b < 2; ++b)
        for (unsigned a = 0; a < 2; ++a) {
          const double value = q[p] * inverseVolume * (a ? fraction[0] : 1. - fraction[0]) *
                               (b ? fraction[1] : 1. - fraction[1]) * (cc ? fraction[2] : 1. - fraction[2]);
          if (!std::isfinite(value)) throw std::invalid_argument("Nonfinite blocked CIC contribution");
          if (value == 0.) continue;
          const std::size_t i = cell[0] + a, j = cell[1] + b, k = cell[2] + cc;
          if (!r.interior(i, j, k))
            throw std::invalid_argument("Blocked source CIC support touches Dirichlet surface; padding unsupported");
          r.rho[((k - r.lower[2] - 1) * r.nodes[1] + (j - r.lower[1] - 1)) * r.nodes[0] + i - r.lower[0] - 1] += value;
        }
  }
  long double deposited = 0.;
  for (double v : r.rho) deposited += static_cast<long double>(v) * volume;
  r.sourceCharge = static_cast<double>(source);
  r.depositedCharge = static_cast<double>(deposited);
}

// depositQuintic, blocked_cowan_reference.cpp:559-593: particles taken by z
// block (stable), each block accumulating the particles of its own and the
// two neighbouring blocks in that order.
inline void depositShaped(Reference& r, std::size_t n, const double* x3, const double* q) {
  const Config& c = r.c;
  if (r.blockLayers < 8) throw std::invalid_argument("Quintic reference requires block_layers >= 8");
  r.rho.assign(r.area * r.nodes[2], 0.);
  std::vector<std::size_t> offsets(r.blocks + 1, 0), order(n);
  long double source = 0.;
  const auto bucket = [&](std::size_t p) {
    const double z = ((x3[3 * p + 2] - c.origin[2]) / c.dz - r.lower[2] - 1) / r.blockLayers;
    if (!std::isfinite(z)) throw std::invalid_argument("Nonfinite quintic reference source");
    return static_cast<std::size_t>(std::max(0., std::min(static_cast<double>(r.blocks - 1), std::floor(z))));
  };
  for (std::size_t p = 0; p < n; ++p) {
    if (!std::isfinite(q[p])) throw std::invalid_argument("Nonfinite quintic reference charge");
    source += q[p];
    ++offsets[bucket(p) + 1];
  }
  for (std::size_t b = 1; b < offsets.size(); ++b) offsets[b] += offsets[b - 1];
  std::vector<std::size_t> next = offsets;
  for (std::size_t p = 0; p < n; ++p) order[next[bucket(p)]++] = p;
  long double deposited = 0.;
  for (std::size_t b = 0; b < r.blocks; ++b) {
    const std::size_t first = b ? b - 1 : 0, last = std::min(r.blocks, b + 2);
    for (std::size_t m = offsets[first]; m < offsets[last]; ++m) {
      const std::size_t p = order[m];
      const double x = (x3[3 * p] - c.origin[0]) / c.dx, y = (x3[3 * p + 1] - c.origin[1]) / c.dy,
                   z = (x3[3 * p + 2] - c.origin[2]) / c.dz;
      const long long ix = static_cast<long long>(std::floor(x)), iy = static_cast<long long>(std::floor(y));
      const double fx = x - ix, fy = y - iy, scale = q[p] / (c.dx * c.dy * c.dz);
      const Basis weights = basis(z, false, c.shape);
      const std::array<double, 4> wx = c.xBinomial ? nodeWeights(fx) : std::array<double, 4>{{1. - fx, fx, 0., 0.}};
      const long long firstX = ix - (c.xBinomial ? 1 : 0);
      const unsigned nX = c.xBinomial ? 4 : 2;
      for (unsigned cc = 0; cc < weights.count; ++cc) {
        const long long k = weights.first + cc;
        if (weights.v[cc] == 0.) continue;
        for (unsigned j = 0; j < 2; ++j)
          for (unsigned i = 0; i < nX; ++i) {
            const double value = scale * wx[i] * (j ? fy : 1. - fy) * weights.v[cc];
            if (value == 0.) continue;
            const long long nodeX = firstX + i, nodeY = iy + j;
            if (k < 0 || nodeX < 0 || nodeY < 0 ||
                !r.interior(static_cast<std::size_t>(nodeX), static_cast<std::size_t>(nodeY),
                            static_cast<std::size_t>(k)))
              throw std::invalid_argument("Reference charge support touches Dirichlet surface");
            const std::size_t kk = static_cast<std::size_t>(k) - r.lower[2] - 1;
            if (kk / r.blockLayers == b)
              r.rho[(kk * r.nodes[1] + (nodeY - r.lower[1] - 1)) * r.nodes[0] + nodeX - r.lower[0] - 1] += value;
          }
      }
    }
    const std::size_t k0 = b * r.blockLayers, k1 = std::min(r.nodes[2], k0 + r.blockLayers);
    for (std::size_t kk = k0; kk < k1; ++kk)
      for (std::size_t m = 0; m < r.area; ++m) {
        const double v = r.rho[kk * r.area + m];
        if (!std::isfinite(v)) throw std::runtime_error("Nonfinite quintic reference density");
        deposited += static_cast<long double>(v) * c.dx * c.dy * c.dz;
      }
  }
  r.sourceCharge = static_cast<double>(source);
  r.depositedCharge = static_cast<double>(deposited);
}

// 2-D DST-I of one interior plane by its definition (FFTW RODFT00 along x
// and y, unnormalized), then scaled by factor.
inline void planeDst(const Reference& r, double* plane, double factor) {
  const std::size_t nx = r.nodes[0], ny = r.nodes[1];
  const double pi = std::acos(-1.);
  std::vector<double> sx(nx * nx), sy(ny * ny), tmp(nx * ny);
  for (std::size_t j = 0; j < nx; ++j)
    for (std::size_t k = 0; k < nx; ++k) sx[j * nx + k] = std::sin(pi * double((j + 1) * (k + 1)) / double(nx + 1));
  for (std::size_t j = 0; j < ny; ++j)
    for (std::size_t k = 0; k < ny; ++k) sy[j * ny + k] = std::sin(pi * double((j + 1) * (k + 1)) / double(ny + 1));
  for (std::size_t y = 0; y < ny; ++y)
    for (std::size_t k = 0; k < nx; ++k) {
      double s = 0.;
      for (std::size_t j = 0; j < nx; ++j) s += plane[y * nx + j] * sx[j * nx + k];
      tmp[y * nx + k] = 2. * s;
    }
  for (std::size_t k = 0; k < ny; ++k)
    for (std::size_t x = 0; x < nx; ++x) {
      double s = 0.;
      for (std::size_t j = 0; j < ny; ++j) s += tmp[j * nx + x] * sy[j * ny + k];
      plane[k * nx + x] = 2. * s * factor;
    }
}

// solve, blocked_cowan_reference.cpp:709-770 (one Thomas sweep per mode over
// every interior plane; blocking only reorders the memory traffic).
inline void solve(Reference& r, const std::vector<double>& diagonal, const std::vector<double>& off) {
  const std::size_t nzn = r.nodes[2], area = r.area;
  std::vector<double> rhs(r.rho);
  for (double& v : rhs) v /= kEps0;
  for (std::size_t k = 0; k < nzn; ++k) planeDst(r, rhs.data() + k * area, 1.);
  std::vector<double> cprime(area * nzn), dprime(area * nzn), cprev(area, 0.), dprev(area, 0.), next(area, 0.);
  r.minimumPivot = std::numeric_limits<double>::infinity();
  for (std::size_t k = 0; k < nzn; ++k)
    for (std::size_t m = 0; m < area; ++m) {
      const double pivot = diagonal[m] - off[m] * cprev[m];
      if (!std::isfinite(pivot) || !(pivot > 0.))
        throw std::runtime_error("Blocked Thomas matrix has nonpositive/nonfinite pivot");
      r.minimumPivot = std::min(r.minimumPivot, pivot);
      cprev[m] = off[m] / pivot;
      dprev[m] = (rhs[k * area + m] - off[m] * dprev[m]) / pivot;
      cprime[k * area + m] = cprev[m];
      dprime[k * area + m] = dprev[m];
    }
  for (std::size_t kk = nzn; kk > 0; --kk)
    for (std::size_t m = 0; m < area; ++m) {
      const std::size_t q = (kk - 1) * area + m;
      dprime[q] -= cprime[q] * next[m];
      next[m] = dprime[q];
    }
  const double normalization = 1. / (4. * (r.nodes[0] + 1) * (r.nodes[1] + 1));
  for (std::size_t k = 0; k < nzn; ++k) planeDst(r, dprime.data() + k * area, normalization);
  r.phi.assign(r.fullArea * (r.c.nz + 1), 0.);
  for (std::size_t k = 0; k < nzn; ++k)
    for (std::size_t j = 0; j < r.nodes[1]; ++j)
      for (std::size_t i = 0; i < r.nodes[0]; ++i)
        r.phiAt(i + r.lower[0] + 1, j + r.lower[1] + 1, k + r.lower[2] + 1) = dprime[(k * r.nodes[1] + j) * r.nodes[0] + i];
}

// smooth, blocked_cowan_reference.cpp:822-841 (phi at global plane k).
inline double smooth(const Reference& r, unsigned axis, std::size_t i, std::size_t j, std::size_t k) {
  const unsigned ta = (axis + 1) % 3, tb = (axis + 2) % 3;
  const std::size_t stride[3] = {1, r.c.nx + 1, r.fullArea};
  const std::size_t q = k * r.fullArea + j * (r.c.nx + 1) + i;
  const std::vector<double>& phi = r.phi;
  double value = r.coefficient.alpha[axis] * phi[q];
  for (int sign = -1; sign <= 1; sign += 2) {
    const std::size_t qa = sign < 0 ? q - stride[ta] : q + stride[ta];
    const std::size_t qb = sign < 0 ? q - stride[tb] : q + stride[tb];
    value += r.coefficient.beta[ta] * phi[qa];
    value += r.coefficient.beta[tb] * phi[qb];
  }
  const double delta = deltaFor(r.coefficient, axis);
  for (int a = -1; a <= 1; a += 2)
    for (int b = -1; b <= 1; b += 2) {
      std::size_t s = a < 0 ? q - stride[ta] : q + stride[ta];
      s = b < 0 ? s - stride[tb] : s + stride[tb];
      value += delta * phi[s];
    }
  return value - (axis == 2 ? r.translation * phi[q] : 0.);
}

// ensureField for the whole grid, blocked_cowan_reference.cpp:862-950: the
// smoothed potentials on interior vertices (zero elsewhere), E as their
// negative differences, B of the rigid translation at -dt/2, Bz = 0.
inline void buildFields(Reference& r) {
  const Config& c = r.c;
  std::array<std::vector<double>, 3> filtered;
  for (auto& f : filtered) f.assign(r.fullArea * (c.nz + 1), 0.);
  for (std::size_t k = r.lower[2] + 1; k < r.upper[2]; ++k)
    for (std::size_t j = r.lower[1] + 1; j < r.upper[1]; ++j)
      for (std::size_t i = r.lower[0] + 1; i < r.upper[0]; ++i)
        for (unsigned axis = 0; axis < 3; ++axis)
          filtered[axis][k * r.fullArea + j * (c.nx + 1) + i] = smooth(r, axis, i, j, k);
  const double spacing[3] = {c.dx, c.dy, c.dz};
  const std::size_t stride[3] = {1, c.nx + 1, r.fullArea};
  for (unsigned comp = 0; comp < 6; ++comp) {
    std::vector<double>& out = r.field[comp];
    out.assign(r.nx[comp] * r.ny[comp] * r.nz[comp], 0.);
    if (comp == 5) continue;
    for (std::size_t k = 0; k < r.nz[comp]; ++k)
      for (std::size_t j = 0; j < r.ny[comp]; ++j)
        for (std::size_t i = 0; i < r.nx[comp]; ++i) {
          const std::size_t q = k * r.fullArea + j * (c.nx + 1) + i;
          double value;
          if (comp < 3) {
            value = -(filtered[comp][q + stride[comp]] - filtered[comp][q]) / spacing[comp];
          } else {
            const unsigned axis = comp == 3 ? 1 : 0;
            const double e0 = -(filtered[axis][q + stride[axis]] - filtered[axis][q]) / spacing[axis];
            const double e1 =
                -(filtered[axis][q + r.fullArea + stride[axis]] - filtered[axis][q + r.fullArea]) / spacing[axis];
            value = (comp == 3 ? -r.magneticScale : r.magneticScale) * (r.lowerWeight * e0 + r.upperWeight * e1);
          }
          out[(k * r.ny[comp] + j) * r.nx[comp] + i] = value;
        }
  }
  for (const auto& f : r.field)
    for (double v : f)
      if (!std::isfinite(v)) throw std::runtime_error("Nonfinite blocked native field");
}

// The whole preparation (prepareCollective on one rank, without files).
inline Reference prepare(const Config& config, std::size_t n, const double* x3, const double* q) {
  Reference r;
  r.c = config;
  configure(r);
  double stencil[3][3][3];
  std::vector<double> diagonal, off;
  makeModes(r, stencil, diagonal, off);
  if (config.shape < 0) depositCic(r, n, x3, q);
  else depositShaped(r, n, x3, q);
  solve(r, diagonal, off);
  buildFields(r);
  return r;
}

// movingWindowFieldValue, simulation_moving_window.cpp:280-294.
inline double movingValue(const Reference& r, unsigned a, std::size_t i, std::size_t j, std::ptrdiff_t k,
                          std::size_t step, double translationBeta) {
  const double z = static_cast<double>(k) - translationBeta * step;
  const std::size_t count = r.c.nz + ((a == 0 || a == 1 || a == 5) ? 1 : 0);
  if (a >= 6 || !std::isfinite(z) || z < 0. || z > count - 1)
    throw std::out_of_range("Moving reference query outside prepared native field");
  const auto lo = static_cast<std::size_t>(std::floor(z)), hi = std::min(lo + 1, count - 1);
  const double f = z - lo, lower = r.value(a, i, j, lo);
  return f == 0. ? lower : (1. - f) * lower + f * r.value(a, i, j, hi);
}

}  // namespace inletref
}  // namespace orc
