// Oracle restatement of the Gauss-consistent initial self-field (TEST
// INFRASTRUCTURE ONLY; never linked into the product):
// EBGaussFieldInitializer::initialize (eb_initial_field.cpp:829-1095) on one
// rank, with its helpers in the same arithmetic order: the CIC charge
// deposit (ChargeConservingCurrentDepositor::depositCharge,
// eb_deposition.cpp:168-193 with normalize / containingCell / clampUnit,
// 248-277), the mean velocity (629-668), the distribution moments
// (698-747), the relativistic-Poisson 7-point and Cowan 27-point operators
// (189-316), conjugate gradients with long-double dot products (1013-1063),
// the fields from the potential (442-575), the CPML clear (578-624), the
// Gauss check (749-811) and the initial field energies (670-696).
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <vector>

#include "orc_fields.hpp"

namespace orc {

enum class InitialModel { RelativisticPoisson = 0, ElectrostaticPoisson = 1, CowanRelativisticPoisson = 2 };

struct GaussReport {
  std::size_t iterations = 0;
  double relativeResidual = 0, maximumGaussResidual = 0, totalCharge = 0, meanBetaZ = 0, meanBetaTransverse = 0,
         initialElectricEnergy = 0, initialMagneticEnergy = 0;
  double rmsPositionCells[3] = {0, 0, 0}, centrePaddingRms[3] = {0, 0, 0};
  std::size_t staticFieldGuardCells[3] = {0, 0, 0}, transversePoissonPaddingCells = 0;
};

struct InitialParticles {
  std::size_t n = 0;
  const double* x = nullptr;  // n x 3
  const double* u = nullptr;  // n x 3 proper velocity / c
  const double* q = nullptr;
  const double* m = nullptr;
};

namespace initial {

struct Domain {
  std::size_t lower[3], upper[3];
};

// Scalar slab of one rank: vertex planes p = 1..nz hold k = p - 1; planes
// 0 and nz + 1 are the zero ghosts (exchangeScalarGhosts without
// neighbours, eb_initial_field.cpp:98-119).
struct Slab {
  std::size_t nx, ny, owned;
  std::vector<double> v;
  Slab(std::size_t a, std::size_t b, std::size_t c) : nx(a), ny(b), owned(c), v(a * b * (c + 2), 0.0) {}
  double& operator()(std::size_t i, std::size_t j, std::size_t p) { return v[(p * ny + j) * nx + i]; }
  double operator()(std::size_t i, std::size_t j, std::size_t p) const { return v[(p * ny + j) * nx + i]; }
  void ghosts() {
    std::fill(v.begin(), v.begin() + nx * ny, 0.0);
    std::fill(v.begin() + (owned + 1) * nx * ny, v.end(), 0.0);
  }
};

inline bool interiorVertex(std::size_t i, std::size_t j, std::size_t k, const Domain& d) {
  return i > d.lower[0] && i < d.upper[0] && j > d.lower[1] && j < d.upper[1] && k > d.lower[2] && k < d.upper[2];
}

// BoostFrameTransform::gammaFromProperVelocity, boostframe.cpp:74-87.
inline double gammaFromU(const double* u) {
  const double scale = std::max(1.0, std::max(std::abs(u[0]), std::max(std::abs(u[1]), std::abs(u[2]))));
  const double inverseScale = 1.0 / scale;
  const double ux = u[0] * inverseScale, uy = u[1] * inverseScale, uz = u[2] * inverseScale;
  const double inverseScale2 = inverseScale * inverseScale;
  return scale * std::sqrt(inverseScale2 + ux * ux + uy * uy + uz * uz);
}

// depositCharge, eb_deposition.cpp:168-193 (with 228-277).
inline void depositCharge(const Grid& g, const double origin[3], const InitialParticles& P, Comp& rho) {
  const double spacing[3] = {g.dx, g.dy, g.dz};
  const std::size_t cells[3] = {g.nx, g.ny, g.nz};
  const double extent[3] = {static_cast<double>(g.nx) * g.dx, static_cast<double>(g.ny) * g.dy,
                            static_cast<double>(g.nz) * g.dz};
  const double inverseVolume = 1.0 / (g.dx * g.dy * g.dz);
  for (std::size_t p = 0; p < P.n; ++p) {
    const double* x = P.x + 3 * p;
    std::size_t cell[3];
    double f[3];
    for (int a = 0; a < 3; ++a) {
      const double relative = x[a] - origin[a];
      const double tolerance = 64.0 * kEpsilon * std::max(1.0, std::max(std::abs(origin[a]), extent[a]));
      if (!std::isfinite(x[a]) || relative < -tolerance || relative > extent[a] + tolerance)
        throw std::out_of_range("Particle trajectory endpoint is outside the E/B domain");
    }
    for (int a = 0; a < 3; ++a) {
      const double c = std::max(0.0, std::min(static_cast<double>(cells[a]), (x[a] - origin[a]) / spacing[a]));
      std::size_t k;
      if (c <= 0.0) k = 0;
      else if (c >= static_cast<double>(cells[a])) k = cells[a] - 1;
      else k = static_cast<std::size_t>(std::floor(c));
      cell[a] = k;
      f[a] = std::max(0.0, std::min(1.0, c - static_cast<double>(k)));
    }
    for (unsigned c = 0; c < 2; ++c)
      for (unsigned b = 0; b < 2; ++b)
        for (unsigned a = 0; a < 2; ++a)
          rho(cell[0] + a, cell[1] + b, cell[2] + c) += P.q[p] * inverseVolume * (a ? f[0] : 1.0 - f[0]) *
                                                        (b ? f[1] : 1.0 - f[1]) * (c ? f[2] : 1.0 - f[2]);
  }
}

// CowanPoissonOperator, eb_initial_field.cpp:209-278.
struct CowanOperator {
  Cowan coefficient;
  double translation = 0.0;
  double stencil[3][3][3] = {};
  CowanOperator(const Grid& g, double meanBetaZ) : coefficient(Cowan::forZDispersion(g.dx, g.dy, g.dz)) {
    const double velocity = meanBetaZ * kC;
    translation = velocity * velocity * kEps0 * kMu0;
    const double spacing[3] = {g.dx, g.dy, g.dz};
    for (unsigned axis = 0; axis < 3; ++axis) {
      const unsigned ta = (axis + 1) % 3, tb = (axis + 2) % 3;
      const double delta = axis == 0 ? coefficient.deltaYZ : (axis == 1 ? coefficient.deltaZX : coefficient.deltaXY);
      double minimum = std::numeric_limits<double>::infinity();
      for (int a = -1; a <= 1; a += 2)
        for (int b = -1; b <= 1; b += 2)
          minimum = std::min(minimum, coefficient.alpha[axis] + 2.0 * coefficient.beta[ta] * a +
                                          2.0 * coefficient.beta[tb] * b + 4.0 * delta * a * b);
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
          const double weight =
              a == 0 ? (b == 0 ? coefficient.alpha[axis] : coefficient.beta[tb]) : (b == 0 ? coefficient.beta[ta] : delta);
          int point[3] = {1, 1, 1};
          point[ta] += a;
          point[tb] += b;
          stencil[point[2]][point[1]][point[0]] += 2.0 * weight * inv2;
          --point[axis];
          stencil[point[2]][point[1]][point[0]] -= weight * inv2;
          point[axis] += 2;
          stencil[point[2]][point[1]][point[0]] -= weight * inv2;
        }
    }
    const double tz = translation / (g.dz * g.dz);
    stencil[1][1][1] -= 2.0 * tz;
    stencil[0][1][1] += tz;
    stencil[2][1][1] += tz;
  }
};

inline void applyLaplacian(const Slab& in, Slab& out, const Grid& g, const Domain& d, double lc) {
  std::fill(out.v.begin(), out.v.end(), 0.0);
  const double cx = 1.0 / (g.dx * g.dx), cy = 1.0 / (g.dy * g.dy), cz = lc / (g.dz * g.dz);
  const double diagonal = 2.0 * (cx + cy + cz);
  for (std::size_t p = 1; p <= in.owned; ++p) {
    const std::size_t k = p - 1;
    if (k <= d.lower[2] || k >= d.upper[2]) continue;
    for (std::size_t j = d.lower[1] + 1; j < d.upper[1]; ++j)
      for (std::size_t i = d.lower[0] + 1; i < d.upper[0]; ++i)
        out(i, j, p) = diagonal * in(i, j, p) - cx * (in(i - 1, j, p) + in(i + 1, j, p)) -
                       cy * (in(i, j - 1, p) + in(i, j + 1, p)) - cz * (in(i, j, p - 1) + in(i, j, p + 1));
  }
}

inline void applyCowan(const Slab& in, Slab& out, const Domain& d, const CowanOperator& op) {
  std::fill(out.v.begin(), out.v.end(), 0.0);
  for (std::size_t p = 1; p <= in.owned; ++p) {
    const std::size_t k = p - 1;
    if (k <= d.lower[2] || k >= d.upper[2]) continue;
    for (std::size_t j = d.lower[1] + 1; j < d.upper[1]; ++j)
      for (std::size_t i = d.lower[0] + 1; i < d.upper[0]; ++i) {
        double value = 0.0;
        for (std::size_t z = 0; z < 3; ++z)
          for (std::size_t y = 0; y < 3; ++y)
            for (std::size_t x = 0; x < 3; ++x) value += op.stencil[z][y][x] * in(i - 1 + x, j - 1 + y, p - 1 + z);
        out(i, j, p) = value;
      }
  }
}

inline double cowanSmoothed(const Slab& s, const Cowan& c, unsigned axis, std::size_t i, std::size_t j,
                            std::size_t p) {
  const unsigned ta = (axis + 1) % 3, tb = (axis + 2) % 3;
  double value = c.alpha[axis] * s(i, j, p);
  for (int sign = -1; sign <= 1; sign += 2) {
    std::size_t a[3] = {i, j, p}, b[3] = {i, j, p};
    if (sign < 0) { --a[ta]; --b[tb]; } else { ++a[ta]; ++b[tb]; }
    value += c.beta[ta] * s(a[0], a[1], a[2]);
    value += c.beta[tb] * s(b[0], b[1], b[2]);
  }
  const double delta = axis == 0 ? c.deltaYZ : (axis == 1 ? c.deltaZX : c.deltaXY);
  for (int a = -1; a <= 1; a += 2)
    for (int b = -1; b <= 1; b += 2) {
      std::size_t point[3] = {i, j, p};
      if (a < 0) --point[ta]; else ++point[ta];
      if (b < 0) --point[tb]; else ++point[tb];
      value += delta * s(point[0], point[1], point[2]);
    }
  return value;
}

inline long double dot(const Slab& l, const Slab& r, const Domain& d) {
  long double sum = 0.0L;
  for (std::size_t p = 1; p <= l.owned; ++p) {
    const std::size_t k = p - 1;
    if (k <= d.lower[2] || k >= d.upper[2]) continue;
    for (std::size_t j = d.lower[1] + 1; j < d.upper[1]; ++j)
      for (std::size_t i = d.lower[0] + 1; i < d.upper[0]; ++i)
        sum += static_cast<long double>(l(i, j, p)) * static_cast<long double>(r(i, j, p));
  }
  return sum;
}

inline void update(Slab& t, const Slab& s, double scale, const Domain& d) {
  for (std::size_t p = 1; p <= t.owned; ++p) {
    const std::size_t k = p - 1;
    if (k <= d.lower[2] || k >= d.upper[2]) continue;
    for (std::size_t j = d.lower[1] + 1; j < d.upper[1]; ++j)
      for (std::size_t i = d.lower[0] + 1; i < d.upper[0]; ++i) t(i, j, p) += scale * s(i, j, p);
  }
}

// B at -dt/2 by rigid translation (eb_initial_field.cpp:475-493, 553-574).
inline void magneticFromElectric(Fields& f, double magneticScale, double meanBetaZ) {
  const Grid& g = f.g;
  const double upper = 0.5 + 0.5 * meanBetaZ * kC * g.dt / g.dz;
  const double lower = 1.0 - upper;
  if (upper < 0.0 || upper > 1.0) throw std::runtime_error("Initial B half-step shift exceeds one longitudinal cell");
  for (std::size_t k = 0; k < g.nz; ++k)
    for (std::size_t j = 0; j < g.ny; ++j)
      for (std::size_t i = 0; i <= g.nx; ++i)
        f.f[3](i, j, k) = -magneticScale * (lower * f.f[1](i, j, k) + upper * f.f[1](i, j, k + 1));
  for (std::size_t k = 0; k < g.nz; ++k)
    for (std::size_t j = 0; j <= g.ny; ++j)
      for (std::size_t i = 0; i < g.nx; ++i)
        f.f[4](i, j, k) = magneticScale * (lower * f.f[0](i, j, k) + upper * f.f[0](i, j, k + 1));
  std::fill(f.f[5].v.begin(), f.f[5].v.end(), 0.0);
}

}  // namespace initial

// EBGaussFieldInitializer::initialize on one rank (no held-transverse mean
// velocity).
inline GaussReport initializeGauss(Fields& f, const double origin[3], const InitialParticles& P, InitialModel model,
                                   double maxMeanTransverseBeta, const std::size_t guard[3], double tolerance,
                                   std::size_t maxIterations, bool includeCpml) {
  using namespace initial;
  if (!(tolerance > 0.0) || !(tolerance < 1.0) || !std::isfinite(tolerance) || maxIterations == 0)
    throw std::invalid_argument("Invalid initial Gauss-field convergence controls");
  if (!(maxMeanTransverseBeta >= 0.0) || !(maxMeanTransverseBeta < 1.0) || !std::isfinite(maxMeanTransverseBeta))
    throw std::invalid_argument("Invalid initial self-field maximum mean transverse beta");
  const Grid& g = f.g;
  const bool cowanModel = model == InitialModel::CowanRelativisticPoisson;
  if (cowanModel && g.solver != kCowanZ)
    throw std::invalid_argument("cowan-relativistic-poisson requires the cowan-z Maxwell solver");
  if (includeCpml && (guard[0] < 2 || guard[1] < 2))
    throw std::invalid_argument(
        "Transverse CPML-source initialization requires at least two physical guard cells in x and y");
  const std::size_t pad = includeCpml ? 2 : 0;
  Grid pg = g;
  pg.nx += 2 * pad;
  pg.ny += 2 * pad;
  double porigin[3] = {origin[0] - static_cast<double>(pad) * g.dx, origin[1] - static_cast<double>(pad) * g.dy,
                       origin[2]};
  Domain domain, physical;
  const std::size_t cells[3] = {g.nx, g.ny, g.nz};
  for (unsigned a = 0; a < 3; ++a) {
    if (2 * guard[a] + 2 >= cells[a])
      throw std::invalid_argument("Initial static-field guard leaves fewer than two physical cells");
    physical.lower[a] = guard[a];
    physical.upper[a] = cells[a] - guard[a];
    domain.lower[a] = includeCpml && a < 2 ? 0 : guard[a];
    const std::size_t pc = a == 0 ? pg.nx : (a == 1 ? pg.ny : pg.nz);
    domain.upper[a] = pc - domain.lower[a];
  }
  for (int a = 0; a < 6; ++a) std::fill(f.f[a].v.begin(), f.f[a].v.end(), 0.0);

  // meanVelocity, 629-668
  long double sums[4] = {};
  for (std::size_t p = 0; p < P.n; ++p) {
    const double gamma = gammaFromU(P.u + 3 * p);
    const long double w = static_cast<long double>(P.m[p]);
    sums[0] += w;
    for (unsigned a = 0; a < 3; ++a) sums[a + 1] += w * static_cast<long double>(P.u[3 * p + a] / gamma);
  }
  if (!(sums[0] > 0.0L)) throw std::runtime_error("Initial relativistic-Poisson field requires positive represented mass");
  const double betaX = static_cast<double>(sums[1] / sums[0]), betaY = static_cast<double>(sums[2] / sums[0]);
  const double betaZ = static_cast<double>(sums[3] / sums[0]);
  const double betaT = std::hypot(betaX, betaY);
  if (!std::isfinite(betaZ) || !std::isfinite(betaT) || betaT >= 1.0 || std::abs(betaZ) >= 1.0)
    throw std::runtime_error("Initial particle mean velocity is not a finite subluminal vector");
  if ((model == InitialModel::RelativisticPoisson || cowanModel) && betaT > maxMeanTransverseBeta) {
    std::ostringstream m;
    m.precision(12);
    m << (cowanModel ? "cowan-relativistic-poisson" : "relativistic-poisson") << " mean beta_box_transverse=" << betaT
      << " exceeds initial_self_field.maximum_mean_transverse_beta=" << maxMeanTransverseBeta
      << "; the rigid-beam initial field assumes a mean velocity along the boost z axis. Align/recenter the beam or "
         "explicitly relax this approximation limit.";
    throw std::runtime_error(m.str());
  }
  const double applied = (model == InitialModel::RelativisticPoisson || cowanModel) ? betaZ : 0.0;
  const double lc = std::fma(-applied, applied, 1.0);
  if (!cowanModel && !(lc > 64.0 * kEpsilon))
    throw std::runtime_error(
        "Initial relativistic-Poisson operator is ill-conditioned because the bunch remains too relativistic in the "
        "simulation frame; choose boost_gamma closer to the bunch mean gamma");
  std::unique_ptr<CowanOperator> op;
  if (cowanModel) op.reset(new CowanOperator(pg, applied));

  Slab potential(pg.nx + 1, pg.ny + 1, g.nz), residual(pg.nx + 1, pg.ny + 1, g.nz),
      direction(pg.nx + 1, pg.ny + 1, g.nz), image(pg.nx + 1, pg.ny + 1, g.nz);
  GaussReport report;
  report.meanBetaZ = betaZ;
  report.meanBetaTransverse = betaT;
  report.transversePoissonPaddingCells = pad;
  for (unsigned a = 0; a < 3; ++a) report.staticFieldGuardCells[a] = domain.lower[a];

  // distributionResolution, 698-747
  {
    long double mom[7] = {};
    for (std::size_t p = 0; p < P.n; ++p) {
      const long double w = static_cast<long double>(P.m[p]);
      mom[0] += w;
      for (unsigned a = 0; a < 3; ++a) {
        const long double x = static_cast<long double>(P.x[3 * p + a]);
        mom[a + 1] += w * x;
        mom[a + 4] += w * x * x;
      }
    }
    if (!(mom[0] > 0.0L)) throw std::runtime_error("Initial field cannot diagnose an empty particle distribution");
    const double spacing[3] = {pg.dx, pg.dy, pg.dz};
    for (unsigned a = 0; a < 3; ++a) {
      const long double mean = mom[a + 1] / mom[0];
      const long double variance = std::max(0.0L, mom[a + 4] / mom[0] - mean * mean);
      const double rms = static_cast<double>(std::sqrt(variance));
      report.rmsPositionCells[a] = rms / spacing[a];
      if (rms > 0.0) {
        const double lower = porigin[a] + static_cast<double>(domain.lower[a]) * spacing[a];
        const double upper = porigin[a] + static_cast<double>(domain.upper[a]) * spacing[a];
        report.centrePaddingRms[a] = std::min(static_cast<double>(mean) - lower, upper - static_cast<double>(mean)) / rms;
      } else {
        report.centrePaddingRms[a] = 0.0;
      }
    }
  }

  Comp rho;
  rho.resize(g.nx + 1, g.ny + 1, g.nz + 1);
  depositCharge(g, origin, P, rho);
  long double total = 0.0L;
  for (std::size_t p = 0; p < P.n; ++p) total += static_cast<long double>(P.q[p]);
  report.totalCharge = static_cast<double>(total);
  long double interior = 0.0L;
  const long double vertexVolume =
      static_cast<long double>(g.dx) * static_cast<long double>(g.dy) * static_cast<long double>(g.dz);
  for (std::size_t p = 1; p <= g.nz; ++p) {
    const std::size_t k = p - 1;
    for (std::size_t j = 0; j <= pg.ny; ++j)
      for (std::size_t i = 0; i <= pg.nx; ++i)
        if (interiorVertex(i, j, k, domain)) {
          const bool inside = i >= pad && i - pad <= g.nx && j >= pad && j - pad <= g.ny;
          const double r = inside ? rho(i - pad, j - pad, k) : 0.0;
          residual(i, j, p) = r / kEps0;
          interior += static_cast<long double>(r) * vertexVolume;
        }
  }
  if (std::abs((interior - total) / total) > 1.0e-12L)
    throw std::runtime_error(
        "Initial CIC charge cloud touches the static zero-potential boundary; move every particle at least one "
        "complete cell inside the CPML entrance or enlarge the physical mesh padding");
  direction = residual;
  const long double source2 = dot(residual, residual, domain);
  if (!(source2 > 0.0L) || !std::isfinite(static_cast<double>(source2)))
    throw std::runtime_error("Initial particle charge produced an invalid Poisson source norm");
  long double r2 = source2;
  for (std::size_t it = 0; it < maxIterations; ++it) {
    if (cowanModel) applyCowan(direction, image, domain, *op);
    else applyLaplacian(direction, image, pg, domain, lc);
    const long double den = dot(direction, image, domain);
    if (!(den > 0.0L) || !std::isfinite(static_cast<double>(den)))
      throw std::runtime_error("Initial Gauss-field CG lost positive definiteness");
    const double alpha = static_cast<double>(r2 / den);
    update(potential, direction, alpha, domain);
    update(residual, image, -alpha, domain);
    const long double next = dot(residual, residual, domain);
    report.iterations = it + 1;
    report.relativeResidual = static_cast<double>(std::sqrt(next / source2));
    if (report.relativeResidual <= tolerance) {
      r2 = next;
      break;
    }
    if (!(next >= 0.0L) || !std::isfinite(static_cast<double>(next)))
      throw std::runtime_error("Initial Gauss-field CG produced a non-finite residual");
    const double beta = static_cast<double>(next / r2);
    for (std::size_t p = 1; p <= direction.owned; ++p) {
      const std::size_t k = p - 1;
      for (std::size_t j = 0; j <= pg.ny; ++j)
        for (std::size_t i = 0; i <= pg.nx; ++i)
          direction(i, j, p) = interiorVertex(i, j, k, domain) ? residual(i, j, p) + beta * direction(i, j, p) : 0.0;
    }
    r2 = next;
  }
  if (report.relativeResidual > tolerance)
    throw std::runtime_error(
        "Initial Gauss-field solve did not converge; increase initial_self_field.maximum_iterations, relax its "
        "relative_tolerance, or reduce the grid aspect ratio");

  if (cowanModel) {
    const double spacing[3] = {g.dx, g.dy, g.dz};
    for (unsigned axis = 0; axis < 3; ++axis) {
      std::fill(image.v.begin(), image.v.end(), 0.0);
      for (std::size_t p = 1; p <= potential.owned; ++p) {
        const std::size_t k = p - 1;
        if (k <= domain.lower[2] || k >= domain.upper[2]) continue;
        for (std::size_t j = domain.lower[1] + 1; j < domain.upper[1]; ++j)
          for (std::size_t i = domain.lower[0] + 1; i < domain.upper[0]; ++i)
            image(i, j, p) = cowanSmoothed(potential, op->coefficient, axis, i, j, p) -
                             (axis == 2 ? op->translation * potential(i, j, p) : 0.0);
      }
      Comp& c = f.f[axis];
      for (std::size_t k = 0; k < c.nz; ++k)
        for (std::size_t j = 0; j < c.ny; ++j)
          for (std::size_t i = 0; i < c.nx; ++i)
            c(i, j, k) = -(image(i + pad + (axis == 0), j + pad + (axis == 1), k + 1 + (axis == 2)) -
                           image(i + pad, j + pad, k + 1)) /
                         spacing[axis];
    }
    magneticFromElectric(f, applied * kC * kEps0 * kMu0, applied);
  } else {
    for (std::size_t k = 0; k <= g.nz; ++k)
      for (std::size_t j = 0; j <= g.ny; ++j)
        for (std::size_t i = 0; i < g.nx; ++i)
          f.f[0](i, j, k) = -(potential(i + pad + 1, j + pad, k + 1) - potential(i + pad, j + pad, k + 1)) / g.dx;
    for (std::size_t k = 0; k <= g.nz; ++k)
      for (std::size_t j = 0; j < g.ny; ++j)
        for (std::size_t i = 0; i <= g.nx; ++i)
          f.f[1](i, j, k) = -(potential(i + pad, j + pad + 1, k + 1) - potential(i + pad, j + pad, k + 1)) / g.dy;
    for (std::size_t k = 0; k < g.nz; ++k)
      for (std::size_t j = 0; j <= g.ny; ++j)
        for (std::size_t i = 0; i <= g.nx; ++i)
          f.f[2](i, j, k) = -lc * (potential(i + pad, j + pad, k + 2) - potential(i + pad, j + pad, k + 1)) / g.dz;
    magneticFromElectric(f, applied / kC, applied);
  }
  if (includeCpml) {
    // clearComponentOutsidePhysical, 578-624
    for (int a = 0; a < 6; ++a) {
      Comp& c = f.f[a];
      const bool electric = a < 3;
      const unsigned ca = static_cast<unsigned>(a % 3);
      for (std::size_t k = 0; k < c.nz; ++k)
        for (std::size_t j = 0; j < c.ny; ++j)
          for (std::size_t i = 0; i < c.nx; ++i) {
            const std::size_t coord[3] = {i, j, k};
            bool inside = true;
            for (unsigned axis = 0; axis < 3; ++axis) {
              const bool vertex = electric ? axis != ca : axis == ca;
              inside = inside && coord[axis] >= guard[axis] &&
                       (vertex ? coord[axis] <= cells[axis] - guard[axis] : coord[axis] < cells[axis] - guard[axis]);
            }
            if (!inside) c(i, j, k) = 0.0;
          }
    }
  }
  // validateGauss, 749-811, with the unpadded charge
  Comp vrho;
  vrho.resize(g.nx + 1, g.ny + 1, g.nz + 1);
  depositCharge(g, origin, P, vrho);
  long double res2 = 0.0L, src2 = 0.0L;
  double maximum = 0.0;
  for (std::size_t k = 0; k < g.nz; ++k) {
    if (k <= physical.lower[2] || k >= physical.upper[2]) continue;
    for (std::size_t j = physical.lower[1] + 1; j < physical.upper[1]; ++j)
      for (std::size_t i = physical.lower[0] + 1; i < physical.upper[0]; ++i) {
        const double leftZ = k > 0 ? f.f[2](i, j, k - 1) : 0.0;
        const double divergence = (f.f[0](i, j, k) - f.f[0](i - 1, j, k)) / g.dx +
                                  (f.f[1](i, j, k) - f.f[1](i, j - 1, k)) / g.dy + (f.f[2](i, j, k) - leftZ) / g.dz;
        const double source = vrho(i, j, k) / kEps0;
        const double error = divergence - source;
        res2 += static_cast<long double>(error) * error;
        src2 += static_cast<long double>(source) * source;
        maximum = std::max(maximum, std::abs(error));
      }
  }
  report.maximumGaussResidual = maximum;
  report.relativeResidual = src2 > 0.0L ? static_cast<double>(std::sqrt(res2 / src2)) : 0.0;
  if (report.relativeResidual > 4.0 * tolerance)
    throw std::runtime_error("Initial E field failed the post-solve discrete Gauss check");
  // initialFieldEnergy, 670-696 (radiationSampleCell, eb_field.cpp:693-713)
  long double el = 0.0L, ma = 0.0L;
  const long double volume = vertexVolume;
  for (std::size_t k = 0; k < g.nz; ++k)
    for (std::size_t j = 0; j < g.ny; ++j)
      for (std::size_t i = 0; i < g.nx; ++i) {
        const Comp *ex = &f.f[0], *ey = &f.f[1], *ez = &f.f[2], *bx = &f.f[3], *by = &f.f[4], *bz = &f.f[5];
        double e[3], b[3];
        e[0] = 0.25 * ((*ex)(i, j, k) + (*ex)(i, j + 1, k) + (*ex)(i, j, k + 1) + (*ex)(i, j + 1, k + 1));
        e[1] = 0.25 * ((*ey)(i, j, k) + (*ey)(i + 1, j, k) + (*ey)(i, j, k + 1) + (*ey)(i + 1, j, k + 1));
        e[2] = 0.25 * ((*ez)(i, j, k) + (*ez)(i + 1, j, k) + (*ez)(i, j + 1, k) + (*ez)(i + 1, j + 1, k));
        b[0] = 0.5 * ((*bx)(i, j, k) + (*bx)(i + 1, j, k));
        b[1] = 0.5 * ((*by)(i, j, k) + (*by)(i, j + 1, k));
        b[2] = 0.5 * ((*bz)(i, j, k) + (*bz)(i, j, k + 1));
        const double e2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
        const double b2 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
        el += 0.5L * static_cast<long double>(kEps0) * static_cast<long double>(e2) * volume;
        ma += 0.5L / static_cast<long double>(kMu0) * static_cast<long double>(b2) * volume;
      }
  report.initialElectricEnergy = static_cast<double>(el);
  report.initialMagneticEnergy = static_cast<double>(ma);
  return report;
}

}  // namespace orc
