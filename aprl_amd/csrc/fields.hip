// gfx950 field kernels: fused Cowan-z (or Yee) magnetic curl + CFS-CPML-B,
// fused electric curl - dt/eps0 J + CFS-CPML-E + PEC + window inlet, and the
// moving-window shift as a base-pointer slide.
//
// Arithmetic follows the reference's association order with contraction
// disabled (-ffp-contract=off), so results are bitwise those of
// eb_field.cpp / eb_cpml.cpp / moving_window_fields.cpp on the same inputs.
#include <cmath>
#include <cstring>
#include <limits>

#include "fields_impl.hpp"

namespace aprl_gpu {

// ---------------------------------------------------------------------------
// Host-side set-up (restated from the reference's constructors).

// EBCowanCoefficients::forZDispersion, eb_field.cpp:33-73.
CowanCoef cowanForZDispersion(double dx, double dy, double dz) {
  if (!(dx > 0.0) || !(dy > 0.0) || !(dz > 0.0) || !std::isfinite(dx) || !std::isfinite(dy) ||
      !std::isfinite(dz))
    throw std::invalid_argument("Cowan-z grid spacing must be positive and finite");
  const double eps = std::numeric_limits<double>::epsilon();
  const double tolerance = 64.0 * eps * std::max(dx, std::max(dy, dz));
  if (dx + tolerance < dz || dy + tolerance < dz)
    throw std::invalid_argument("Cowan-z requires dz to be the smallest grid spacing");
  CowanCoef r{};
  r.ratio[0] = (dz / dx) * (dz / dx);
  r.ratio[1] = (dz / dy) * (dz / dy);
  r.ratio[2] = 1.0;
  const double rx = r.ratio[0], ry = r.ratio[1], rz = r.ratio[2];
  const double den = rx * ry + ry * rz + rz * rx;
  if (!(den > 0.0) || !std::isfinite(den)) throw std::invalid_argument("Invalid Cowan-z aspect ratios");
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

// EBFieldGrid::validateGeometry, eb_field.cpp:878-943.
void validateGeometry(const aprl_grid_geometry& g, const CowanCoef& cw) {
  const double eps = std::numeric_limits<double>::epsilon();
  if (g.nx < 2 || g.ny < 2 || g.nz < 2)
    throw std::invalid_argument("E/B staggered grid needs at least two cells in every direction");
  if (!(g.dx > 0.0) || !(g.dy > 0.0) || !(g.dz > 0.0) || !(g.dt > 0.0) || !std::isfinite(g.dx) ||
      !std::isfinite(g.dy) || !std::isfinite(g.dz) || !std::isfinite(g.dt))
    throw std::invalid_argument("E/B grid spacing and time step must be positive finite SI values");
  if (g.nx + 1 > (1u << 30) || g.ny + 1 > (1u << 30) || g.nz + 1 > (1u << 30))
    throw std::invalid_argument("E/B grid extent exceeds the device index range");
  if (g.solver == 0) {
    const double courant = kC * g.dt * std::sqrt(1.0 / (g.dx * g.dx) + 1.0 / (g.dy * g.dy) + 1.0 / (g.dz * g.dz));
    if (courant > 1.0 + 16.0 * eps)
      throw std::invalid_argument("E/B time step violates the three-dimensional Yee CFL limit");
    return;
  }
  const double tol = 64.0 * eps * std::max(g.dx, std::max(g.dy, g.dz));
  if (g.dx + tol < g.dz || g.dy + tol < g.dz)
    throw std::invalid_argument("Cowan-z requires dz to be the smallest grid spacing");
  if (std::abs(kC * g.dt / g.dz - 1.0) > 64.0 * eps)
    throw std::invalid_argument("Cowan-z requires the dispersion-matched step c*dt=dz");
  const double norm[3] = {cw.alpha[0] + 2.0 * cw.beta[1] + 2.0 * cw.beta[2] + 4.0 * cw.deltaYZ,
                          cw.alpha[1] + 2.0 * cw.beta[2] + 2.0 * cw.beta[0] + 4.0 * cw.deltaZX,
                          cw.alpha[2] + 2.0 * cw.beta[0] + 2.0 * cw.beta[1] + 4.0 * cw.deltaXY};
  for (double v : norm)
    if (!std::isfinite(v) || std::abs(v - 1.0) > 1.0e-12)
      throw std::invalid_argument("Cowan-z smoothing coefficients are not normalized");
  const double rx = cw.ratio[0], ry = cw.ratio[1], rz = cw.ratio[2];
  const double fac[4] = {4.0 * ((cw.beta[2] + 2.0 * cw.deltaZX) / rz + (cw.beta[1] + 2.0 * cw.deltaXY) / ry),
                         4.0 * ((cw.beta[0] + 2.0 * cw.deltaXY) / rx + (cw.beta[2] + 2.0 * cw.deltaYZ) / rz),
                         4.0 * ((cw.beta[1] + 2.0 * cw.deltaYZ) / ry + (cw.beta[0] + 2.0 * cw.deltaZX) / rx),
                         16.0 * (cw.deltaYZ / (ry * rz) + cw.deltaZX / (rz * rx) + cw.deltaXY / (rx * ry))};
  for (double v : fac)
    if (!std::isfinite(v) || std::abs(v - 1.0) > 1.0e-12)
      throw std::invalid_argument("Cowan-z coefficients do not satisfy the stable dispersion factorization");
}

APRL_FIELDS_INSTANTIATE(double)

}  // namespace aprl_gpu
