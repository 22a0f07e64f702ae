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
// Device memory one shaped-charge deposit works in: acc holds one 64-bit
// fixed-point word per vertex and must be zero on entry (every pass clears
// what it reads); partial holds 2 x 512 doubles, out 8 doubles, maxima 4
// words.
struct RhoWorkspace {
  unsigned long long* acc = nullptr;
  double* partial = nullptr;
  double* out = nullptr;
  unsigned long long* maxima = nullptr;
};
// The workspace of s.diag, grown to the grid (allocates on first use).
template <class R>
RhoWorkspace diagnosticWorkspace(FieldStateT<R>& s);
// Shaped vertex charge density (shape 0 quintic, 2 binomial3-x) of
// particles at absolute positions x, or (split form) at lattice cell +
// offset with the grid origin at lattice cell oc, into rho (one double per
// vertex).
void chargeDensity(FieldState& s, int shape, const double* const x[3], const double* q, int n, const double origin[3],
                   const ChargeWeights& cw, const RhoWorkspace& ws, double* rho);
void chargeDensitySplit(FieldState32& s, int shape, const int* const cell[3], const float* const off[3],
                        const float* q, int n, const int oc[3], const ChargeWeights& cw, const RhoWorkspace& ws,
                        double* rho);
// out = {max |(rho1 - rho0)/dt + div J|, max |(rho1 - rho0)/dt|,
// max max(|rho0|, |rho1|)/dt}.
template <class R>
void continuityResidual(FieldStateT<R>& s, const double* rho0, const double* rho1, const RhoWorkspace& ws,
                        double out[3]);
// Vertices of the allocation-free continuity workspace, and that workspace
// with its two charge-density arrays (this device's copy of a file-scope
// device array; throws std::invalid_argument beyond the limit). The
// accumulator is cleared by a fill kernel on s's stream.
std::size_t noAllocVertexLimit();
template <class R>
RhoWorkspace noAllocWorkspace(FieldStateT<R>& s, std::size_t vertices, double** rho0, double** rho1);
// out = {sum error^2, sum source^2, max |error|, max |source|} over the
// vertices strictly inside [lo, hi], error = div E - rho / eps0.
template <class R>
void gaussResidual(FieldStateT<R>& s, const DeviceArray<double>& rho, const std::size_t lo[3], const std::size_t hi[3],
                   double out[4]);

// Moving-window energy ledger (moving_window_fields.cpp:207-218, 264-288,
// 387-438, 443-517). Enable allocates and clears the accumulators; begin
// records the step's start power; commitElectric replaces the window
// commit's fused inlet overwrite by the E update, the overwrite and its
// plane-energy / inlet-power measurement; finish adds the trapezoid face
// flux; beforeShift / afterShift measure the removed and inserted planes.
template <class R>
void ledgerEnable(FieldStateT<R>& s, bool enable);
template <class R>
void ledgerBegin(FieldStateT<R>& s);
template <class R>
void ledgerCommitElectric(FieldStateT<R>& s, const FieldPtrsT<R>& f);
template <class R>
void ledgerFinish(FieldStateT<R>& s);
template <class R>
void ledgerBeforeShift(FieldStateT<R>& s);
template <class R>
void ledgerAfterShift(FieldStateT<R>& s);
// out[0..3] inlet overwrite, inserted, cropped, remap energy; out[4..9]
// outward energy per face; out[10..12] overwrite events, flux steps, remap
// steps.
template <class R>
void ledgerRead(FieldStateT<R>& s, double out[13]);
// Sums over the store's particles (active and held entries) at box time
// `time`: out[0] kinetic, out[1] total energy, out[2] represented
// electrons, out[3..5] represented-weighted lab gamma, box gamma and lab z,
// out[6] lab kinetic energy, out[7] / out[8] c p_z in the box / lab.
template <class R, class Store>
void particleEnergySums(FieldStateT<R>& s, const struct CouplingState& C, const Store& P, double time,
                        double out[9]);
// The window commit: fused E update with the inlet overwrite, or with the
// ledger on, ledgerCommitElectric.
template <class R>
void windowCommitElectric(FieldStateT<R>& s, const FieldPtrsT<R>& f);

}  // namespace aprl_gpu
