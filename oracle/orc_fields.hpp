// Oracle: CPU restatement of APRL's field step (TEST INFRASTRUCTURE ONLY).
//
// This file is the checker, never the product: only tests/, the smoke entry
// point and the bench CPU-baseline leg may load it. It restates, in plain
// serial C++, the arithmetic of the reference files cited at each function,
// in the same association order, so that it can be compared bitwise with a
// build of the reference's own translation units (oracle/_ref) and, within
// tolerance or bitwise, with the HIP library.
//
// Pinning of every function here (see tests/test_oracle_*.py):
//   * Cowan coefficients: closed form + normalization/factorization identities
//     + the equal-spacing values alpha=7/12, beta=1/12, delta=1/48.
//   * Cowan B / E update / PEC / CPML / window boundary / shift: bitwise
//     equality with oracle/_ref on seeded fields, plus the axial plane-wave
//     exact-translation test (core_checks.cpp:771-823) and the CPML
//     reflection bound (core_checks.cpp:952-1030).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <vector>

namespace orc {

// SI constants, eb_field.cpp:15-19.
constexpr double kC = 299792458.0;
constexpr double kEps0 = 8.8541878128e-12;
constexpr double kMu0 = 1.25663706212e-6;
constexpr double kQe = 1.602176634e-19;
constexpr double kMe = 9.1093837139e-31;
constexpr double kEpsilon = std::numeric_limits<double>::epsilon();

enum Solver { kYee = 0, kCowanZ = 1 };

struct Grid {
  std::size_t nx = 0, ny = 0, nz = 0;
  double dx = 0, dy = 0, dz = 0, dt = 0;
  int solver = kYee;
};

// EBCowanCoefficients::forZDispersion, eb_field.cpp:33-73.
struct Cowan {
  double alpha[3] = {1, 1, 1};
  double beta[3] = {0, 0, 0};
  double deltaXY = 0, deltaYZ = 0, deltaZX = 0;
  double ratio[3] = {1, 1, 1};

  static Cowan forZDispersion(double dx, double dy, double dz) {
    if (!(dx > 0.0) || !(dy > 0.0) || !(dz > 0.0) || !std::isfinite(dx) ||
        !std::isfinite(dy) || !std::isfinite(dz))
      throw std::invalid_argument("Cowan-z grid spacing must be positive and finite");
    const double tolerance = 64.0 * kEpsilon * std::max(dx, std::max(dy, dz));
    if (dx + tolerance < dz || dy + tolerance < dz)
      throw std::invalid_argument("Cowan-z requires dz to be the smallest grid spacing");
    Cowan r;
    r.ratio[0] = (dz / dx) * (dz / dx);
    r.ratio[1] = (dz / dy) * (dz / dy);
    r.ratio[2] = 1.0;
    const double rx = r.ratio[0], ry = r.ratio[1], rz = r.ratio[2];
    const double den = rx * ry + ry * rz + rz * rx;
    if (!(den > 0.0) || !std::isfinite(den))
      throw std::invalid_argument("Invalid Cowan-z aspect ratios");
    const double pf = rx * ry * rz / den;
    r.beta[0] = rx * (1.0 - pf) / 8.0;
    r.beta[1] = ry * (1.0 - pf) / 8.0;
    r.beta[2] = rz * (1.0 - pf) / 8.0;
    r.deltaXY = rx * ry * (1.0 / 16.0 - rx * ry / (8.0 * den));
    r.deltaYZ = ry * rz * (1.0 / 16.0 - ry * rz / (8.0 * den));
    r.deltaZX = rz * rx * (1.0 / 16.0 - rz * rx / (8.0 * den));
    r.alpha[0] = 1.0 - 2.0 * r.beta[1] - 2.0 * r.beta[2] - 4.0 * r.deltaYZ;
    r.alpha[1] = 1.0 - 2.0 * r.beta[2] - 2.0 * r.beta[0] - 4.0 * r.deltaZX;
    r.alpha[2] = 1.0 - 2.0 * r.beta[0] - 2.0 * r.beta[1] - 4.0 * r.deltaXY;
    return r;
  }
};

// YeeComponent storage, eb_field.h:108-147: x fastest, (k*ny+j)*nx+i.
struct Comp {
  std::size_t nx = 0, ny = 0, nz = 0;
  std::vector<double> v;
  void resize(std::size_t a, std::size_t b, std::size_t c) {
    nx = a; ny = b; nz = c; v.assign(a * b * c, 0.0);
  }
  double& operator()(std::size_t i, std::size_t j, std::size_t k) { return v[(k * ny + j) * nx + i]; }
  double operator()(std::size_t i, std::size_t j, std::size_t k) const { return v[(k * ny + j) * nx + i]; }
  std::size_t size() const { return v.size(); }
  double* data() { return v.data(); }
  const double* data() const { return v.data(); }
};

// Component shapes, eb_field.cpp:217-225 (E: +1 off-axis; B: +1 on-axis).
inline void componentShape(const Grid& g, unsigned a, std::size_t s[3]) {
  const bool electric = a < 3 || a >= 6;
  const unsigned axis = a % 3;
  const std::size_t cells[3] = {g.nx, g.ny, g.nz};
  for (unsigned d = 0; d < 3; ++d)
    s[d] = cells[d] + (electric ? (d != axis ? 1 : 0) : (d == axis ? 1 : 0));
}

struct Halo {  // EBElectricHaloView, eb_field.h:76-86; null = PEC parity.
  const double* lower[3] = {nullptr, nullptr, nullptr};
  const double* upper[3] = {nullptr, nullptr, nullptr};
};

struct CpmlParams {  // EBCPMLParameters, eb_cpml.h:12-24 / eb_cpml.cpp:251-256.
  std::size_t cells[3] = {0, 0, 0};
  double order = 3.0, reflection = 1.0e-8, kappaMax = 8.0, alphaFraction = 0.0;
  bool enabled() const { return cells[0] > 0 || cells[1] > 0 || cells[2] > 0; }
  // eb_cpml.cpp:263-281.
  double maxRate(unsigned axis, const Grid& g) const {
    if (cells[axis] == 0) return 0.0;
    const double spacing = axis == 0 ? g.dx : (axis == 1 ? g.dy : g.dz);
    if (!(spacing > 0.0) || !std::isfinite(spacing) || !(order > 0.0) ||
        !(reflection > 0.0 && reflection < 1.0))
      throw std::invalid_argument("Cannot evaluate an invalid CPML conductivity profile");
    const double thickness = static_cast<double>(cells[axis]) * spacing;
    return -(order + 1.0) * kC * std::log(reflection) / (2.0 * thickness);
  }
};

struct CpmlPoint { double ikm1 = 0, decay = 0, drive = 0; bool active = false; };
struct CpmlProfile { unsigned axis = 0; std::vector<CpmlPoint> point; std::vector<std::size_t> active; };

// CPMLMemoryTerm, eb_cpml.cpp:42-96.
struct CpmlTerm {
  unsigned axis = 0;
  std::size_t dim[3] = {0, 0, 0};
  std::vector<std::size_t> base;
  std::vector<double> hist;
  void init(unsigned a, const std::size_t shape[3], const CpmlProfile& p) {
    axis = a;
    for (unsigned d = 0; d < 3; ++d) dim[d] = shape[d];
    const std::size_t missing = std::numeric_limits<std::size_t>::max();
    base.assign(dim[axis], missing);
    const std::size_t cross = axis == 0 ? dim[1] * dim[2] : (axis == 1 ? dim[0] * dim[2] : dim[0] * dim[1]);
    hist.assign(p.active.size() * cross, 0.0);
    for (std::size_t o = 0; o < p.active.size(); ++o) base[p.active[o]] = o * cross;
  }
  double& value(std::size_t i, std::size_t j, std::size_t k) {
    const std::size_t ai = axis == 0 ? i : (axis == 1 ? j : k);
    const std::size_t b = base[ai];
    if (b == std::numeric_limits<std::size_t>::max()) throw std::logic_error("Inactive CPML history was requested");
    const std::size_t ci = axis == 0 ? k * dim[1] + j : (axis == 1 ? k * dim[0] + i : j * dim[0] + i);
    return hist[b + ci];
  }
};

// makeProfile, eb_cpml.cpp:195-248.
inline CpmlProfile makeProfile(unsigned axis, bool electric, const Grid& g, std::size_t globalNz,
                               std::size_t zOffset, const CpmlParams& p) {
  CpmlProfile r;
  r.axis = axis;
  const std::size_t localCells = axis == 0 ? g.nx : (axis == 1 ? g.ny : g.nz);
  const std::size_t globalCells = axis == 2 ? globalNz : localCells;
  const std::size_t globalOffset = axis == 2 ? zOffset : 0;
  const std::size_t count = localCells + (electric ? 1 : 0);
  r.point.resize(count);
  const std::size_t thicknessCells = p.cells[axis];
  if (thicknessCells == 0) return r;
  const double thickness = static_cast<double>(thicknessCells);
  const double smax = p.maxRate(axis, g);
  for (std::size_t li = 0; li < count; ++li) {
    const double position = static_cast<double>(globalOffset + li) + (electric ? 0.0 : 0.5);
    double depth = 0.0;
    if (position < thickness) depth = (thickness - position) / thickness;
    const double upperStart = static_cast<double>(globalCells - thicknessCells);
    if (position > upperStart) depth = std::max(depth, (position - upperStart) / thickness);
    if (!(depth > 0.0)) continue;
    depth = std::min(1.0, depth);
    const double graded = std::pow(depth, p.order);
    const double sigma = smax * graded;
    const double kappa = 1.0 + (p.kappaMax - 1.0) * graded;
    const double alpha = p.alphaFraction * smax * (1.0 - depth);
    CpmlPoint& c = r.point[li];
    c.ikm1 = 1.0 / kappa - 1.0;
    c.decay = std::exp(-(sigma / kappa + alpha) * g.dt);
    const double den = sigma * kappa + alpha * kappa * kappa;
    c.drive = den > 0.0 ? sigma * (c.decay - 1.0) / den : 0.0;
    c.active = true;
    r.active.push_back(li);
  }
  return r;
}

// transformedCorrection, eb_cpml.cpp:131-138.
inline double transformedCorrection(const CpmlPoint& c, double raw, double& history) {
  history = c.decay * history + c.drive * raw;
  return c.ikm1 * raw + history;
}

struct Cpml {
  CpmlParams p;
  bool lowerZ = true, upperZ = true;
  CpmlProfile mag[3], ele[3];
  CpmlTerm magTerm[3][3], eleTerm[3][3];
  bool hasMag[3][3] = {}, hasEle[3][3] = {};
};

constexpr unsigned kPairs[6][2] = {{0, 1}, {0, 2}, {1, 2}, {1, 0}, {2, 0}, {2, 1}};

struct Window;

// EBFieldGrid restatement, eb_field.h:201-357 / eb_field.cpp.
class Fields {
 public:
  Grid g;
  Cowan cw;
  Comp f[9];  // Ex Ey Ez Bx By Bz Jx Jy Jz
  std::unique_ptr<Cpml> cpml;  // null => PerfectElectricConductorBoundary
  bool pecLowerZ = true, pecUpperZ = true;
  std::unique_ptr<Window> window;

  explicit Fields(const Grid& grid);
  ~Fields();

  // eb_field.cpp:878-943.
  void validate() const {
    if (g.nx < 2 || g.ny < 2 || g.nz < 2)
      throw std::invalid_argument("E/B staggered grid needs at least two cells in every direction");
    if (!(g.dx > 0.0) || !(g.dy > 0.0) || !(g.dz > 0.0) || !(g.dt > 0.0) || !std::isfinite(g.dx) ||
        !std::isfinite(g.dy) || !std::isfinite(g.dz) || !std::isfinite(g.dt))
      throw std::invalid_argument("E/B grid spacing and time step must be positive finite SI values");
    if (g.solver == kYee) {
      const double courant = kC * g.dt * std::sqrt(1.0 / (g.dx * g.dx) + 1.0 / (g.dy * g.dy) + 1.0 / (g.dz * g.dz));
      if (courant > 1.0 + 16.0 * kEpsilon)
        throw std::invalid_argument("E/B time step violates the three-dimensional Yee CFL limit");
      return;
    }
    const double tol = 64.0 * kEpsilon * std::max(g.dx, std::max(g.dy, g.dz));
    if (g.dx + tol < g.dz || g.dy + tol < g.dz)
      throw std::invalid_argument("Cowan-z requires dz to be the smallest grid spacing");
    const double axial = kC * g.dt / g.dz;
    if (std::abs(axial - 1.0) > 64.0 * kEpsilon)
      throw std::invalid_argument("Cowan-z requires the dispersion-matched step c*dt=dz");
    const double norm[3] = {cw.alpha[0] + 2.0 * cw.beta[1] + 2.0 * cw.beta[2] + 4.0 * cw.deltaYZ,
                            cw.alpha[1] + 2.0 * cw.beta[2] + 2.0 * cw.beta[0] + 4.0 * cw.deltaZX,
                            cw.alpha[2] + 2.0 * cw.beta[0] + 2.0 * cw.beta[1] + 4.0 * cw.deltaXY};
    for (unsigned a = 0; a < 3; ++a)
      if (!std::isfinite(norm[a]) || std::abs(norm[a] - 1.0) > 1.0e-12)
        throw std::invalid_argument("Cowan-z smoothing coefficients are not normalized");
    const double rx = cw.ratio[0], ry = cw.ratio[1], rz = cw.ratio[2];
    const double fac[4] = {4.0 * ((cw.beta[2] + 2.0 * cw.deltaZX) / rz + (cw.beta[1] + 2.0 * cw.deltaXY) / ry),
                           4.0 * ((cw.beta[0] + 2.0 * cw.deltaXY) / rx + (cw.beta[2] + 2.0 * cw.deltaYZ) / rz),
                           4.0 * ((cw.beta[1] + 2.0 * cw.deltaYZ) / ry + (cw.beta[0] + 2.0 * cw.deltaZX) / rx),
                           16.0 * (cw.deltaYZ / (ry * rz) + cw.deltaZX / (rz * rx) + cw.deltaXY / (rx * ry))};
    for (unsigned c = 0; c < 4; ++c)
      if (!std::isfinite(fac[c]) || std::abs(fac[c] - 1.0) > 1.0e-12)
        throw std::invalid_argument("Cowan-z coefficients do not satisfy the stable dispersion factorization");
  }

  // electricValue, eb_field.cpp:350-417. Component a in 0..2.
  double electricValue(unsigned a, std::ptrdiff_t i, std::ptrdiff_t j, std::ptrdiff_t k, const Halo& h) const {
    const Comp& c = f[a];
    std::ptrdiff_t x[3] = {i, j, k};
    const std::ptrdiff_t n[3] = {(std::ptrdiff_t)c.nx, (std::ptrdiff_t)c.ny, (std::ptrdiff_t)c.nz};
    double sign = 1.0;
    for (unsigned ax = 0; ax < 2; ++ax) {
      if (x[ax] < 0) {
        if (a == ax) x[ax] = 0;
        else { x[ax] = -x[ax]; sign = -sign; }
      } else if (x[ax] >= n[ax]) {
        if (a == ax) x[ax] = n[ax] - 1;
        else { x[ax] = 2 * (n[ax] - 1) - x[ax]; sign = -sign; }
      }
    }
    if (x[2] < 0 || x[2] >= n[2]) {
      const bool lower = x[2] < 0;
      const double* plane = lower ? h.lower[a] : h.upper[a];
      if (plane) return sign * plane[(std::size_t)x[1] * c.nx + (std::size_t)x[0]];
      if (a == 2) x[2] = lower ? 0 : n[2] - 1;
      else { x[2] = lower ? -x[2] : 2 * (n[2] - 1) - x[2]; sign = -sign; }
    }
    return sign * c((std::size_t)x[0], (std::size_t)x[1], (std::size_t)x[2]);
  }

  // smoothedElectric, eb_field.cpp:419-486. The interior fast path has the
  // same accumulation order as the general path, so one path suffices.
  double smoothedElectric(unsigned a, unsigned d, std::ptrdiff_t i, std::ptrdiff_t j, std::ptrdiff_t k,
                          const Halo& h) const {
    const unsigned tA = (d + 1) % 3, tB = (d + 2) % 3;
    const double delta = d == 0 ? cw.deltaYZ : (d == 1 ? cw.deltaZX : cw.deltaXY);
    const std::ptrdiff_t base[3] = {i, j, k};
    double r = cw.alpha[d] * electricValue(a, i, j, k, h);
    for (int dir = -1; dir <= 1; dir += 2) {
      std::ptrdiff_t pA[3] = {base[0], base[1], base[2]};
      std::ptrdiff_t pB[3] = {base[0], base[1], base[2]};
      pA[tA] += dir;
      pB[tB] += dir;
      r += cw.beta[tA] * electricValue(a, pA[0], pA[1], pA[2], h);
      r += cw.beta[tB] * electricValue(a, pB[0], pB[1], pB[2], h);
    }
    for (int dA = -1; dA <= 1; dA += 2)
      for (int dB = -1; dB <= 1; dB += 2) {
        std::ptrdiff_t p[3] = {base[0], base[1], base[2]};
        p[tA] += dA;
        p[tB] += dB;
        r += delta * electricValue(a, p[0], p[1], p[2], h);
      }
    return r;
  }

  // advanceMagneticYee, eb_field.cpp:311-348.
  void advanceMagneticYee() {
    const double tx = g.dt / g.dx, ty = g.dt / g.dy, tz = g.dt / g.dz;
    Comp &ex = f[0], &ey = f[1], &ez = f[2], &bx = f[3], &by = f[4], &bz = f[5];
    for (std::size_t k = 0; k < g.nz; ++k)
      for (std::size_t j = 0; j < g.ny; ++j)
        for (std::size_t i = 0; i <= g.nx; ++i)
          bx(i, j, k) -= ty * (ez(i, j + 1, k) - ez(i, j, k)) - tz * (ey(i, j, k + 1) - ey(i, j, k));
    for (std::size_t k = 0; k < g.nz; ++k)
      for (std::size_t j = 0; j <= g.ny; ++j)
        for (std::size_t i = 0; i < g.nx; ++i)
          by(i, j, k) -= tz * (ex(i, j, k + 1) - ex(i, j, k)) - tx * (ez(i + 1, j, k) - ez(i, j, k));
    for (std::size_t k = 0; k <= g.nz; ++k)
      for (std::size_t j = 0; j < g.ny; ++j)
        for (std::size_t i = 0; i < g.nx; ++i)
          bz(i, j, k) -= tx * (ey(i + 1, j, k) - ey(i, j, k)) - ty * (ex(i, j + 1, k) - ex(i, j, k));
  }

  // advanceMagneticCowan, eb_field.cpp:488-558.
  void advanceMagneticCowan(const Halo& h) {
    const double tx = g.dt / g.dx, ty = g.dt / g.dy, tz = g.dt / g.dz;
    Comp &bx = f[3], &by = f[4], &bz = f[5];
    for (std::size_t k = 0; k < g.nz; ++k)
      for (std::size_t j = 0; j < g.ny; ++j)
        for (std::size_t i = 0; i <= g.nx; ++i) {
          const double dEzDy = smoothedElectric(2, 1, i, j + 1, k, h) - smoothedElectric(2, 1, i, j, k, h);
          const double dEyDz = smoothedElectric(1, 2, i, j, k + 1, h) - smoothedElectric(1, 2, i, j, k, h);
          bx(i, j, k) -= ty * dEzDy - tz * dEyDz;
        }
    for (std::size_t k = 0; k < g.nz; ++k)
      for (std::size_t j = 0; j <= g.ny; ++j)
        for (std::size_t i = 0; i < g.nx; ++i) {
          const double dExDz = smoothedElectric(0, 2, i, j, k + 1, h) - smoothedElectric(0, 2, i, j, k, h);
          const double dEzDx = smoothedElectric(2, 0, i + 1, j, k, h) - smoothedElectric(2, 0, i, j, k, h);
          by(i, j, k) -= tz * dExDz - tx * dEzDx;
        }
    for (std::size_t k = 0; k <= g.nz; ++k)
      for (std::size_t j = 0; j < g.ny; ++j)
        for (std::size_t i = 0; i < g.nx; ++i) {
          const double dEyDx = smoothedElectric(1, 0, i + 1, j, k, h) - smoothedElectric(1, 0, i, j, k, h);
          const double dExDy = smoothedElectric(0, 1, i, j + 1, k, h) - smoothedElectric(0, 1, i, j, k, h);
          bz(i, j, k) -= tx * dEyDx - ty * dExDy;
        }
  }

  // advanceElectric interior curl, eb_field.cpp:582-624 (boundary not applied).
  void electricCurl() {
    const double curlScale = g.dt / (kMu0 * kEps0);
    const double cs = g.dt / kEps0;
    const double cX = curlScale / g.dx, cY = curlScale / g.dy, cZ = curlScale / g.dz;
    Comp &ex = f[0], &ey = f[1], &ez = f[2], &bx = f[3], &by = f[4], &bz = f[5], &jx = f[6], &jy = f[7], &jz = f[8];
    for (std::size_t k = 1; k < g.nz; ++k)
      for (std::size_t j = 1; j < g.ny; ++j)
        for (std::size_t i = 0; i < g.nx; ++i)
          ex(i, j, k) += cY * (bz(i, j, k) - bz(i, j - 1, k)) - cZ * (by(i, j, k) - by(i, j, k - 1)) - cs * jx(i, j, k);
    for (std::size_t k = 1; k < g.nz; ++k)
      for (std::size_t j = 0; j < g.ny; ++j)
        for (std::size_t i = 1; i < g.nx; ++i)
          ey(i, j, k) += cZ * (bx(i, j, k) - bx(i, j, k - 1)) - cX * (bz(i, j, k) - bz(i - 1, j, k)) - cs * jy(i, j, k);
    for (std::size_t k = 0; k < g.nz; ++k)
      for (std::size_t j = 1; j < g.ny; ++j)
        for (std::size_t i = 1; i < g.nx; ++i)
          ez(i, j, k) += cX * (by(i, j, k) - by(i - 1, j, k)) - cY * (bx(i, j, k) - bx(i, j - 1, k)) - cs * jz(i, j, k);
  }

  // PerfectElectricConductorBoundary::afterElectricUpdate, eb_field.cpp:158-204.
  void pecTerminate(bool lowerZ, bool upperZ) {
    Comp &ex = f[0], &ey = f[1], &ez = f[2];
    for (std::size_t k = 0; k <= g.nz; ++k)
      for (std::size_t i = 0; i < g.nx; ++i) { ex(i, 0, k) = 0.0; ex(i, g.ny, k) = 0.0; }
    if (lowerZ || upperZ)
      for (std::size_t j = 0; j <= g.ny; ++j)
        for (std::size_t i = 0; i < g.nx; ++i) {
          if (lowerZ) ex(i, j, 0) = 0.0;
          if (upperZ) ex(i, j, g.nz) = 0.0;
        }
    for (std::size_t k = 0; k <= g.nz; ++k)
      for (std::size_t j = 0; j < g.ny; ++j) { ey(0, j, k) = 0.0; ey(g.nx, j, k) = 0.0; }
    if (lowerZ || upperZ)
      for (std::size_t j = 0; j < g.ny; ++j)
        for (std::size_t i = 0; i <= g.nx; ++i) {
          if (lowerZ) ey(i, j, 0) = 0.0;
          if (upperZ) ey(i, j, g.nz) = 0.0;
        }
    for (std::size_t k = 0; k < g.nz; ++k)
      for (std::size_t j = 0; j <= g.ny; ++j) { ez(0, j, k) = 0.0; ez(g.nx, j, k) = 0.0; }
    for (std::size_t k = 0; k < g.nz; ++k)
      for (std::size_t i = 0; i <= g.nx; ++i) { ez(i, 0, k) = 0.0; ez(i, g.ny, k) = 0.0; }
  }

  // EBConvolutionalPML construction, eb_cpml.cpp:297-372.
  void installCpml(const CpmlParams& p, std::size_t globalNz, std::size_t zOffset, bool lowerZ, bool upperZ) {
    if (!p.enabled()) throw std::invalid_argument("CPML requires at least one nonzero boundary thickness");
    if (!(p.order > 0.0) || !(p.reflection > 0.0 && p.reflection < 1.0) || !(p.kappaMax >= 1.0) ||
        !(p.alphaFraction >= 0.0) || !std::isfinite(p.order) || !std::isfinite(p.reflection) ||
        !std::isfinite(p.kappaMax) || !std::isfinite(p.alphaFraction))
      throw std::invalid_argument("Invalid CPML grading parameters");
    const std::size_t gc[3] = {g.nx, g.ny, globalNz};
    for (unsigned a = 0; a < 3; ++a)
      if (p.cells[a] > 0) {
        if (p.cells[a] < 2) throw std::invalid_argument("Each enabled CPML direction needs at least two cells");
        if (2 * p.cells[a] >= gc[a]) throw std::invalid_argument("Opposite CPML slabs must leave a non-PML interior");
      }
    const std::size_t zThickness = p.cells[2];
    if (zThickness > 0) {
      if (!lowerZ && zOffset < zThickness)
        throw std::invalid_argument("The lower z CPML slab cannot span MPI ranks yet; reduce MPI ranks or CPML cells");
      if (!upperZ && zOffset + g.nz > globalNz - zThickness)
        throw std::invalid_argument("The upper z CPML slab cannot span MPI ranks yet; reduce MPI ranks or CPML cells");
    }
    std::unique_ptr<Cpml> c(new Cpml);
    c->p = p;
    c->lowerZ = lowerZ;
    c->upperZ = upperZ;
    for (unsigned a = 0; a < 3; ++a) {
      c->mag[a] = makeProfile(a, false, g, globalNz, zOffset, p);
      c->ele[a] = makeProfile(a, true, g, globalNz, zOffset, p);
    }
    for (unsigned pr = 0; pr < 6; ++pr) {
      const unsigned comp = kPairs[pr][0], der = kPairs[pr][1];
      std::size_t s[3];
      componentShape(g, comp + 3, s);
      c->magTerm[comp][der].init(der, s, c->mag[der]);
      c->hasMag[comp][der] = true;
      componentShape(g, comp, s);
      c->eleTerm[comp][der].init(der, s, c->ele[der]);
      c->hasEle[comp][der] = true;
    }
    cpml = std::move(c);
    pecLowerZ = lowerZ;
    pecUpperZ = upperZ;
  }

  // checkpointHistory order, eb_cpml.cpp:546-577.
  std::vector<double>& history(std::size_t index) {
    const unsigned comp = kPairs[index % 6][0], der = kPairs[index % 6][1];
    return index < 6 ? cpml->magTerm[comp][der].hist : cpml->eleTerm[comp][der].hist;
  }

  // EBConvolutionalPML::afterMagneticUpdate, eb_cpml.cpp:406-468.
  void cpmlMagnetic(const Halo& h) {
    const double scale[3] = {g.dt / g.dx, g.dt / g.dy, g.dt / g.dz};
    auto term = [&](unsigned tc, unsigned sc, unsigned d, double sign) {
      CpmlProfile& prof = cpml->mag[d];
      if (prof.active.empty()) return;
      CpmlTerm& t = cpml->magTerm[tc][d];
      Comp& target = f[tc + 3];
      const Comp& source = f[sc];
      forActive(prof, 0, 0, 0, target.nx, target.ny, target.nz, [&](std::size_t i, std::size_t j, std::size_t k) {
        std::ptrdiff_t plus[3] = {(std::ptrdiff_t)i, (std::ptrdiff_t)j, (std::ptrdiff_t)k};
        ++plus[d];
        double diff;
        if (g.solver == kCowanZ)
          diff = smoothedElectric(sc, d, plus[0], plus[1], plus[2], h) -
                 smoothedElectric(sc, d, (std::ptrdiff_t)i, (std::ptrdiff_t)j, (std::ptrdiff_t)k, h);
        else
          diff = d == 0 ? source(i + 1, j, k) - source(i, j, k)
                        : (d == 1 ? source(i, j + 1, k) - source(i, j, k) : source(i, j, k + 1) - source(i, j, k));
        double& hist = t.value(i, j, k);
        const double corr = transformedCorrection(prof.point[d == 0 ? i : (d == 1 ? j : k)], diff, hist);
        target(i, j, k) += sign * scale[d] * corr;
      });
    };
    term(0, 2, 1, -1.0);
    term(0, 1, 2, 1.0);
    term(1, 0, 2, -1.0);
    term(1, 2, 0, 1.0);
    term(2, 1, 0, -1.0);
    term(2, 0, 1, 1.0);
  }

  // EBConvolutionalPML::afterElectricUpdate, eb_cpml.cpp:470-539.
  void cpmlElectric() {
    const double curlScale = g.dt / (kMu0 * kEps0);
    const double scale[3] = {curlScale / g.dx, curlScale / g.dy, curlScale / g.dz};
    auto term = [&](unsigned tc, unsigned sc, unsigned d, double sign) {
      CpmlProfile& prof = cpml->ele[d];
      if (prof.active.empty()) return;
      CpmlTerm& t = cpml->eleTerm[tc][d];
      Comp& target = f[tc];
      const Comp& source = f[sc + 3];
      std::size_t b[3] = {0, 0, 0}, e[3] = {target.nx, target.ny, target.nz};
      if (tc == 0) {
        b[1] = 1; e[1] = g.ny;
        b[2] = cpml->lowerZ ? 1 : 0; e[2] = cpml->upperZ ? g.nz : g.nz + 1;
      } else if (tc == 1) {
        b[0] = 1; e[0] = g.nx;
        b[2] = cpml->lowerZ ? 1 : 0; e[2] = cpml->upperZ ? g.nz : g.nz + 1;
      } else {
        b[0] = 1; e[0] = g.nx; b[1] = 1; e[1] = g.ny;
      }
      forActive(prof, b[0], b[1], b[2], e[0], e[1], e[2], [&](std::size_t i, std::size_t j, std::size_t k) {
        const double diff = d == 0 ? source(i, j, k) - source(i - 1, j, k)
                                   : (d == 1 ? source(i, j, k) - source(i, j - 1, k) : source(i, j, k) - source(i, j, k - 1));
        double& hist = t.value(i, j, k);
        const double corr = transformedCorrection(prof.point[d == 0 ? i : (d == 1 ? j : k)], diff, hist);
        target(i, j, k) += sign * scale[d] * corr;
      });
    };
    term(0, 2, 1, 1.0);
    term(0, 1, 2, -1.0);
    term(1, 0, 2, 1.0);
    term(1, 2, 0, -1.0);
    term(2, 1, 0, 1.0);
    term(2, 0, 1, -1.0);
    pecTerminate(cpml->lowerZ, cpml->upperZ);
  }

  // forActivePoints, eb_cpml.cpp:140-193 (serial; order within a term does
  // not affect results because each point is written once per term).
  template <class Op>
  static void forActive(const CpmlProfile& p, std::size_t b0, std::size_t b1, std::size_t b2, std::size_t e0,
                        std::size_t e1, std::size_t e2, Op op) {
    if (p.axis == 0) {
      for (std::size_t k = b2; k < e2; ++k)
        for (std::size_t j = b1; j < e1; ++j)
          for (std::size_t o = 0; o < p.active.size(); ++o) {
            const std::size_t i = p.active[o];
            if (i >= b0 && i < e0) op(i, j, k);
          }
    } else if (p.axis == 1) {
      for (std::size_t k = b2; k < e2; ++k)
        for (std::size_t o = 0; o < p.active.size(); ++o) {
          const std::size_t j = p.active[o];
          if (j < b1 || j >= e1) continue;
          for (std::size_t i = b0; i < e0; ++i) op(i, j, k);
        }
    } else {
      for (std::size_t o = 0; o < p.active.size(); ++o)
        for (std::size_t j = b1; j < e1; ++j) {
          const std::size_t k = p.active[o];
          if (k < b2 || k >= e2) continue;
          for (std::size_t i = b0; i < e0; ++i) op(i, j, k);
        }
    }
  }

  // EBFieldGrid::advanceMagnetic, eb_field.cpp:302-309 (+ boundary hook).
  void advanceMagnetic(const Halo& h);
  // EBFieldGrid::advanceElectric, eb_field.cpp:582-627 (+ boundary hook).
  void advanceElectric();
  void afterElectricBoundary();
  void installWindow(const CpmlParams& p);
  Halo beginWindowStep(std::size_t n, std::ptrdiff_t k, const double* const lowerPlanes[3],
                       const double* ezAtOffset = nullptr);
  void setWindowInlet(const double* ex, const double* ey);
  void undoPreview();
  void shiftWindow(std::size_t n, std::ptrdiff_t k, const double* const planes[6]);
  // MovingWindowFields energy ledger: planeEnergy, boundaryPower,
  // finishEnergyLedgerStep (moving_window_fields.cpp:387-438) and the
  // shift's inserted-plane measurement (503-509).
  long double planeEnergy(std::size_t k) const;
  void boundaryPower(bool inletOnly, long double power[6]) const;
  void finishEnergyLedgerStep();
  void ledgerInserted();

  void clearCurrent() {
    for (unsigned a = 6; a < 9; ++a) std::fill(f[a].v.begin(), f[a].v.end(), 0.0);
  }

  // fieldEnergyDensityCell, eb_field.cpp:715-738.
  double energyDensityCell(std::size_t i, std::size_t j, std::size_t k) const {
    const Comp &ex = f[0], &ey = f[1], &ez = f[2], &bx = f[3], &by = f[4], &bz = f[5];
    double e0 = 0.25 * (ex(i, j, k) + ex(i, j + 1, k) + ex(i, j, k + 1) + ex(i, j + 1, k + 1));
    double e1 = 0.25 * (ey(i, j, k) + ey(i + 1, j, k) + ey(i, j, k + 1) + ey(i + 1, j, k + 1));
    double e2 = 0.25 * (ez(i, j, k) + ez(i + 1, j, k) + ez(i, j + 1, k) + ez(i + 1, j + 1, k));
    double b0 = 0.5 * (bx(i, j, k) + bx(i + 1, j, k));
    double b1 = 0.5 * (by(i, j, k) + by(i, j + 1, k));
    double b2 = 0.5 * (bz(i, j, k) + bz(i, j, k + 1));
    return 0.5 * (kEps0 * (e0 * e0 + e1 * e1 + e2 * e2) + (b0 * b0 + b1 * b1 + b2 * b2) / kMu0);
  }

  // poyntingComponentCell, eb_field.cpp:740-782.
  double poyntingCell(std::size_t i, std::size_t j, std::size_t k, unsigned axis) const {
    const Comp &ex = f[0], &ey = f[1], &ez = f[2], &bx = f[3], &by = f[4], &bz = f[5];
    auto cEx = [&]() { return 0.25 * (ex(i, j, k) + ex(i, j + 1, k) + ex(i, j, k + 1) + ex(i, j + 1, k + 1)); };
    auto cEy = [&]() { return 0.25 * (ey(i, j, k) + ey(i + 1, j, k) + ey(i, j, k + 1) + ey(i + 1, j, k + 1)); };
    auto cEz = [&]() { return 0.25 * (ez(i, j, k) + ez(i + 1, j, k) + ez(i, j + 1, k) + ez(i + 1, j + 1, k)); };
    auto cBx = [&]() { return 0.5 * (bx(i, j, k) + bx(i + 1, j, k)); };
    auto cBy = [&]() { return 0.5 * (by(i, j, k) + by(i, j + 1, k)); };
    auto cBz = [&]() { return 0.5 * (bz(i, j, k) + bz(i, j, k + 1)); };
    double eA, eB, bA, bB;
    if (axis == 0) { eA = cEy(); eB = cEz(); bA = cBy(); bB = cBz(); }
    else if (axis == 1) { eA = cEz(); eB = cEx(); bA = cBz(); bB = cBx(); }
    else { eA = cEx(); eB = cEy(); bA = cBx(); bB = cBy(); }
    double s = eA * bB - eB * bA;
    s /= kMu0;
    return s;
  }
};

// MovingWindowFields single-rank restatement: Impl::Boundary
// (moving_window_fields.cpp:150-299) and Impl::shift (439-518).
// Reference values are supplied as host planes per call instead of a
// callback; the caller queries the same (a,i,j,k,step) points.
struct Window {
  std::vector<double> lower[3], upper[3];
  std::size_t step = 0;
  std::ptrdiff_t offset = 0;
  bool established = false;
  int phase = 0;  // 0 Idle, 1 Magnetic, 2 Electric, 3 Complete
  bool electricPreview = false, previewPresent = false;
  struct Delta { unsigned axis; std::size_t index; double value; };
  std::vector<Delta> delta;
  std::vector<double> inlet[2];  // Ex/Ey(k=0) := reference(a,i,j,offset,step+1)
  bool inletSet = false;
  // setNewLayerCompletion(..., centerNormalReference = true)
  // (moving_window_fields.cpp:551-555): the inlet Ez is completed from the
  // charge and its departure from the reference is reflected oddly.
  bool centered = false;
  // MovingWindowFieldEnergyLedger (moving_window_fields.h:39-58) and the
  // boundary's per-step ledger state (moving_window_fields.cpp:159-161).
  struct Ledger {
    bool enabled = false;
    long double inletOverwriteEnergyJ = 0., insertedEnergyJ = 0., croppedEnergyJ = 0., remapEnergyJ = 0.;
    long double outwardEnergyFace[6] = {0., 0., 0., 0., 0., 0.};
    std::uint64_t overwriteEvents = 0, fluxSteps = 0, remapSteps = 0;
    long double startPower[6] = {0., 0., 0., 0., 0., 0.};
    long double overwritePowerDelta[6] = {0., 0., 0., 0., 0., 0.};
    bool energyStepFinished = false;
    // a charge-matched window measures the inserted plane after its inlet
    // completion (the reference runs the completion inside the shift)
    bool insertPending = false;
    long double removedLocal = 0.;
  } ledger;
};

inline Fields::Fields(const Grid& grid) : g(grid) {
  cw = g.solver == kCowanZ ? Cowan::forZDispersion(g.dx, g.dy, g.dz) : Cowan();
  for (unsigned a = 0; a < 9; ++a) {
    std::size_t s[3];
    componentShape(g, a, s);
    f[a].resize(s[0], s[1], s[2]);
  }
  validate();
}
inline Fields::~Fields() {}

inline void Fields::advanceMagnetic(const Halo& h) {
  if (window && window->phase != 1) throw std::logic_error("Moving-window magnetic phase mismatch");
  if (g.solver == kCowanZ) advanceMagneticCowan(h);
  else advanceMagneticYee();
  if (cpml) cpmlMagnetic(h);
  if (window) window->phase = 2;
}

inline void Fields::advanceElectric() {
  electricCurl();
  afterElectricBoundary();
}

inline void Fields::afterElectricBoundary() {
  if (!window) {
    if (cpml) cpmlElectric();
    else pecTerminate(pecLowerZ, pecUpperZ);
    return;
  }
  Window& w = *window;
  if (w.phase != 2) throw std::logic_error("Moving-window electric phase mismatch");
  if (!w.inletSet) throw std::logic_error("Moving-window inlet planes were not supplied");
  if (w.electricPreview) {
    // previewElectric, moving_window_fields.cpp:168-186.
    if (w.previewPresent) throw std::logic_error("Unconsumed electric boundary preview");
    w.delta.clear();
    const std::size_t cx = cpml->p.cells[0], cy = cpml->p.cells[1];
    for (unsigned a = 0; a < 3; ++a) {
      const Comp& v = f[a];
      for (std::size_t k = 0; k < v.nz; ++k)
        for (std::size_t j = 0; j < v.ny; ++j)
          for (std::size_t i = 0; i < v.nx; ++i)
            if (i <= cx + 2 || i + cx + 3 >= v.nx || j <= cy + 2 || j + cy + 3 >= v.ny || (a < 2 && k == 0))
              w.delta.push_back({a, (k * v.ny + j) * v.nx + i, v(i, j, k)});
    }
    std::vector<std::vector<double>> saved(6);
    for (std::size_t hh = 6; hh < 12; ++hh) saved[hh - 6] = history(hh);
    cpmlElectric();
    for (unsigned a = 0; a < 2; ++a) {
      Comp& c = f[a];
      for (std::size_t j = 0; j < c.ny; ++j)
        for (std::size_t i = 0; i < c.nx; ++i) c(i, j, 0) = w.inlet[a][j * c.nx + i];
    }
    for (auto& q : w.delta) q.value = f[q.axis].v[q.index] - q.value;
    for (std::size_t hh = 6; hh < 12; ++hh) history(hh).swap(saved[hh - 6]);
    w.previewPresent = true;
    return;
  }
  // Boundary::afterElectricUpdate, moving_window_fields.cpp:254-293.
  cpmlElectric();
  long double energyBefore = 0.;
  long double powerBefore[6];
  const bool measure = w.ledger.enabled;
  if (measure) {
    energyBefore = planeEnergy(0);
    boundaryPower(true, powerBefore);
  }
  for (unsigned a = 0; a < 2; ++a) {
    Comp& c = f[a];
    for (std::size_t j = 0; j < c.ny; ++j)
      for (std::size_t i = 0; i < c.nx; ++i) c(i, j, 0) = w.inlet[a][j * c.nx + i];
  }
  if (measure) {
    w.ledger.inletOverwriteEnergyJ += planeEnergy(0) - energyBefore;
    long double powerAfter[6];
    boundaryPower(true, powerAfter);
    for (unsigned face = 0; face < 6; ++face) w.ledger.overwritePowerDelta[face] = powerAfter[face] - powerBefore[face];
    ++w.ledger.overwriteEvents;
  }
  w.phase = 3;
}

// planeEnergy, moving_window_fields.cpp:387-397 (physical aperture).
inline long double Fields::planeEnergy(std::size_t k) const {
  const long double volume = static_cast<long double>(g.dx) * static_cast<long double>(g.dy) *
                             static_cast<long double>(g.dz);
  const std::size_t cx = cpml->p.cells[0], cy = cpml->p.cells[1];
  long double value = 0.;
  for (std::size_t j = cy; j < g.ny - cy; ++j)
    for (std::size_t i = cx; i < g.nx - cx; ++i) value += static_cast<long double>(energyDensityCell(i, j, k)) * volume;
  if (!std::isfinite(value)) throw std::runtime_error("Non-finite moving-window energy diagnostic");
  return value;
}

// boundaryPower, moving_window_fields.cpp:398-419 (one rank: z- and z+).
inline void Fields::boundaryPower(bool inletOnly, long double power[6]) const {
  for (unsigned face = 0; face < 6; ++face) power[face] = 0.;
  const std::size_t i0 = cpml->p.cells[0], i1 = g.nx - i0;
  const std::size_t j0 = cpml->p.cells[1], j1 = g.ny - j0;
  const long double yz = static_cast<long double>(g.dy) * g.dz;
  const long double xz = static_cast<long double>(g.dx) * g.dz;
  const long double xy = static_cast<long double>(g.dx) * g.dy;
  for (std::size_t k = 0; k < (inletOnly ? 1 : g.nz); ++k) {
    for (std::size_t j = j0; j < j1; ++j) {
      power[0] -= static_cast<long double>(poyntingCell(i0, j, k, 0)) * yz;
      power[1] += static_cast<long double>(poyntingCell(i1 - 1, j, k, 0)) * yz;
    }
    for (std::size_t i = i0; i < i1; ++i) {
      power[2] -= static_cast<long double>(poyntingCell(i, j0, k, 1)) * xz;
      power[3] += static_cast<long double>(poyntingCell(i, j1 - 1, k, 1)) * xz;
    }
  }
  for (std::size_t j = j0; j < j1; ++j)
    for (std::size_t i = i0; i < i1; ++i) {
      power[4] -= static_cast<long double>(poyntingCell(i, j, 0, 2)) * xy;
      if (!inletOnly) power[5] += static_cast<long double>(poyntingCell(i, j, g.nz - 1, 2)) * xy;
    }
  for (unsigned face = 0; face < 6; ++face)
    if (!std::isfinite(power[face])) throw std::runtime_error("Non-finite moving-window boundary power");
}

// finishEnergyLedgerStep, moving_window_fields.cpp:425-438.
inline void Fields::finishEnergyLedgerStep() {
  Window& w = *window;
  if (!w.ledger.enabled) return;
  if (w.phase != 3 || w.ledger.energyStepFinished)
    throw std::logic_error("Moving-window energy flux requires one completed MPI field step");
  long double endPower[6];
  boundaryPower(false, endPower);
  const long double halfDt = .5L * static_cast<long double>(g.dt);
  for (unsigned face = 0; face < 6; ++face)
    w.ledger.outwardEnergyFace[face] +=
        halfDt * (w.ledger.startPower[face] + endPower[face] - w.ledger.overwritePowerDelta[face]);
  ++w.ledger.fluxSteps;
  w.ledger.energyStepFinished = true;
}

// The shift's remap bookkeeping after the new plane is complete,
// moving_window_fields.cpp:503-509.
inline void Fields::ledgerInserted() {
  Window& w = *window;
  const long double insertedLocal = planeEnergy(0);
  w.ledger.remapEnergyJ += insertedLocal - w.ledger.removedLocal;
  w.ledger.insertedEnergyJ += insertedLocal;
  w.ledger.croppedEnergyJ += w.ledger.removedLocal;
  ++w.ledger.remapSteps;
  w.ledger.insertPending = false;
}

// MovingWindowFields construction + installBoundary,
// moving_window_fields.cpp:329-362, 193-200, 526-531 (single rank).
inline void Fields::installWindow(const CpmlParams& p) {
  if (g.solver != kCowanZ || p.cells[2] != 0 || p.cells[0] < 2 || p.cells[1] < 2 || p.cells[0] >= g.nx ||
      p.cells[0] >= g.nx - p.cells[0] || p.cells[1] >= g.ny || p.cells[1] >= g.ny - p.cells[1])
    throw std::invalid_argument("Moving-window requires valid Cowan integer z slabs, xy-only CPML and an inlet reference");
  if (!std::isfinite(g.dt) || !std::isfinite(g.dz) || !(g.dz > 0) ||
      std::abs(kC * g.dt - g.dz) > 64 * kEpsilon * g.dz)
    throw std::invalid_argument("Moving-window requires c*dt=dz");
  installCpml(p, g.nz, 0, false, false);
  window.reset(new Window);
  for (unsigned a = 0; a < 3; ++a) {
    window->lower[a].assign(f[a].nx * f[a].ny, 0.0);
    window->upper[a].assign(f[a].nx * f[a].ny, 0.0);
  }
  for (unsigned a = 0; a < 2; ++a) window->inlet[a].assign(f[a].nx * f[a].ny, 0.0);
}

// Boundary::begin + halos, moving_window_fields.cpp:207-244.
// lowerPlanes[a] = reference(a,i,j,offset-1,step); with a centered normal
// reference, ezAtOffset = reference(2,i,j,offset,step) and the Ez halo of
// the aperture nodes takes minus the inlet Ez's departure from it (225-234).
inline Halo Fields::beginWindowStep(std::size_t n, std::ptrdiff_t k, const double* const lowerPlanes[3],
                                    const double* ezAtOffset) {
  Window& w = *window;
  if (w.phase != 0 || (w.established && (n != w.step || k != w.offset)))
    throw std::logic_error("Moving-window beginStep phase/origin mismatch");
  if (w.centered && !ezAtOffset) throw std::logic_error("A centered inlet needs the reference Ez at the offset");
  if (w.ledger.insertPending) ledgerInserted();
  w.step = n; w.offset = k; w.established = true;
  if (w.ledger.enabled) {  // moving_window_fields.cpp:214-218
    boundaryPower(false, w.ledger.startPower);
    for (unsigned face = 0; face < 6; ++face) w.ledger.overwritePowerDelta[face] = 0.;
    w.ledger.energyStepFinished = false;
  }
  for (unsigned a = 0; a < 3; ++a)
    std::memcpy(w.lower[a].data(), lowerPlanes[a], w.lower[a].size() * sizeof(double));
  if (w.centered) {
    const Comp& ez = f[2];
    const std::size_t cx = cpml->p.cells[0], cy = cpml->p.cells[1];
    for (std::size_t j = cy + 1; j < g.ny - cy; ++j)
      for (std::size_t i = cx + 1; i < g.nx - cx; ++i)
        w.lower[2][j * ez.nx + i] -= ez(i, j, 0) - ezAtOffset[j * ez.nx + i];
  }
  w.phase = 1;
  w.inletSet = false;
  Halo h;
  for (unsigned a = 0; a < 3; ++a) { h.lower[a] = w.lower[a].data(); h.upper[a] = w.upper[a].data(); }
  return h;
}

// inlet planes: Ex/Ey reference(a,i,j,offset,step+1).
inline void Fields::setWindowInlet(const double* ex, const double* ey) {
  Window& w = *window;
  std::memcpy(w.inlet[0].data(), ex, w.inlet[0].size() * sizeof(double));
  std::memcpy(w.inlet[1].data(), ey, w.inlet[1].size() * sizeof(double));
  w.inletSet = true;
}

// Boundary::undoPreview, moving_window_fields.cpp:187-191.
inline void Fields::undoPreview() {
  Window& w = *window;
  if (!w.previewPresent) throw std::logic_error("Missing electric boundary preview");
  for (const auto& q : w.delta) f[q.axis].v[q.index] -= q.value;
  w.previewPresent = false;
}

// Impl::shift with resetCurrent=true, moving_window_fields.cpp:439-518.
// planes[a] = reference(a,i,j,newOffset,newStep) for a in 0..5.
inline void Fields::shiftWindow(std::size_t n, std::ptrdiff_t k, const double* const planes[6]) {
  Window& w = *window;
  if (w.phase != 3 || n != w.step + 1 || k != w.offset - 1)
    throw std::logic_error("Moving-window requires one left shift after each complete step");
  if (w.ledger.enabled && !w.ledger.energyStepFinished)
    throw std::logic_error("Moving-window energy flux must finish after MPI and before shift");
  if (w.ledger.enabled) w.ledger.removedLocal = planeEnergy(g.nz - 1);
  for (unsigned a = 0; a < 6; ++a) {
    Comp& c = f[a];
    const std::size_t plane = c.nx * c.ny;
    std::memmove(c.v.data() + plane, c.v.data(), (c.nz - 1) * plane * sizeof(double));
    std::memcpy(c.v.data(), planes[a], plane * sizeof(double));
  }
  for (unsigned a = 6; a < 9; ++a) std::fill(f[a].v.begin(), f[a].v.end(), 0.0);
  // 12 histories: [slot][layer][row]; memmove layers up by one, insert zeros.
  for (std::size_t h = 0; h < 12; ++h) {
    std::vector<double>& values = history(h);
    if (values.empty()) continue;
    const unsigned comp = kPairs[h % 6][0], der = kPairs[h % 6][1];
    std::size_t s[3];
    componentShape(g, h < 6 ? comp + 3 : comp, s);
    const std::size_t layers = s[2], row = s[der == 0 ? 1 : 0];
    const std::size_t slots = values.size() / (layers * row);
    for (std::size_t slot = 0; slot < slots; ++slot) {
      double* base = values.data() + slot * layers * row;
      std::memmove(base + row, base, (layers - 1) * row * sizeof(double));
      std::fill(base, base + row, 0.0);
    }
  }
  w.step = n; w.offset = k; w.phase = 0;
  if (w.ledger.enabled) {
    w.ledger.insertPending = true;
    if (!w.centered) ledgerInserted();
  }
}

}  // namespace orc
