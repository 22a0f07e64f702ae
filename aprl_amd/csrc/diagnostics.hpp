// FP64 field and coupling diagnostics (energy ledger, continuity, Gauss).
#pragma once

#include "fields.hpp"

namespace aprl_gpu {

// Sum of fieldEnergyDensityCell * dx dy dz over the cells of [lo, hi).
double interiorEnergy(FieldState& s, const std::size_t lo[3], const std::size_t hi[3]);
// Outward Poynting power through the six faces of the cell box [lo, hi):
// power[2a] = -sum over the lower face, power[2a+1] = +sum over the upper.
void facePower(FieldState& s, const std::size_t lo[3], const std::size_t hi[3], double power[6]);
// Shaped vertex charge density (shape 0 quintic, 2 binomial3-x).
void chargeDensity(FieldState& s, int shape, const double* const x[3], const double* q, int n, const double origin[3],
                   DeviceArray<double>& rho);
// out = {max |(rho1 - rho0)/dt + div J|, max |(rho1 - rho0)/dt|}.
void continuityResidual(FieldState& s, const DeviceArray<double>& rho0, const DeviceArray<double>& rho1,
                        double out[2]);
// out = {sum error^2, sum source^2, max |error|, max |source|} over the
// vertices strictly inside [lo, hi], error = div E - rho / eps0.
void gaussResidual(FieldState& s, const DeviceArray<double>& rho, const std::size_t lo[3], const std::size_t hi[3],
                   double out[4]);

}  // namespace aprl_gpu
