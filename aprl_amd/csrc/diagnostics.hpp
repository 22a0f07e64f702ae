// FP64 field and coupling diagnostics (energy ledger, continuity, Gauss).
#pragma once

#include "fields.hpp"

namespace aprl_gpu {

// All sums run in FP64 on either form of the grid (R = double or float).

// Sum of fieldEnergyDensityCell * dx dy dz over the cells of [lo, hi).
template <class R>
double interiorEnergy(FieldStateT<R>& s, const std::size_t lo[3], const std::size_t hi[3]);
// Outward Poynting power through the six faces of the cell box [lo, hi):
// power[2a] = -sum over the lower face, power[2a+1] = +sum over the upper.
template <class R>
void facePower(FieldStateT<R>& s, const std::size_t lo[3], const std::size_t hi[3], double power[6]);
// Charge factor of the store entries: mode 0 all 1; mode 1 the current
// entries (removed ones 0, carriers their current weight w); mode 2 the
// entries at the start of the last step (event code evt: 0 means an
// unchanged particle, otherwise w holds the start weight).
struct ChargeWeights {
  int mode = 0;
  const int* state = nullptr;
  const int* evt = nullptr;
  const double* wd = nullptr;  // weights of the plain form, or
  const float* wf = nullptr;   // of the split form
};
// Shaped vertex charge density (shape 0 quintic, 2 binomial3-x) of
// particles at absolute positions x, or (split form) at lattice cell +
// offset with the grid origin at lattice cell oc.
void chargeDensity(FieldState& s, int shape, const double* const x[3], const double* q, int n, const double origin[3],
                   const ChargeWeights& cw, DeviceArray<double>& rho);
void chargeDensitySplit(FieldState32& s, int shape, const int* const cell[3], const float* const off[3],
                        const float* q, int n, const int oc[3], const ChargeWeights& cw, DeviceArray<double>& rho);
// out = {max |(rho1 - rho0)/dt + div J|, max |(rho1 - rho0)/dt|}.
template <class R>
void continuityResidual(FieldStateT<R>& s, const DeviceArray<double>& rho0, const DeviceArray<double>& rho1,
                        double out[2]);
// out = {sum error^2, sum source^2, max |error|, max |source|} over the
// vertices strictly inside [lo, hi], error = div E - rho / eps0.
template <class R>
void gaussResidual(FieldStateT<R>& s, const DeviceArray<double>& rho, const std::size_t lo[3], const std::size_t hi[3],
                   double out[4]);

}  // namespace aprl_gpu
