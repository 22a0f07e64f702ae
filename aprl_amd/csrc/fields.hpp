// Device-resident Yee field state and its host-side bookkeeping, for the
// plain FP64 form (R = double) and the FP32 form (R = float).
//
// E and B components carry `slack` spare z planes in front of the window so
// that the moving-window shift (moving_window_fields.cpp:482-498) becomes a
// base-pointer decrement plus one inserted plane; a full recentering copy
// happens once every `slack` steps. CPML histories with a transverse
// derivative axis get the same per-slot slack on their z layers.
#pragma once

#include "common.hpp"

struct aprl_inlet_reference;

namespace aprl_gpu {

struct CowanCoef {
  double alpha[3], beta[3], deltaXY, deltaYZ, deltaZX, ratio[3];
};
CowanCoef cowanForZDispersion(double dx, double dy, double dz);

// Profiles and coefficients are FP64 for both storage types R: updates
// of the FP32 form are evaluated in FP64 and rounded once on store.
template <class R>
struct CpmlPointT { double ikm1, decay, drive; };

// Kernel-side view of the field arrays (base pointers already slid).
template <class R>
struct FieldPtrsT {
  R* e[3];
  R* b[3];
  R* j[3];
  int n[9][3];     // (nx, ny, nz) per component
  const R* lo[3];  // lower E halo planes (null => PEC parity)
  const R* up[3];  // upper E halo planes
};

template <class R>
struct StepCoefsT {
  double alpha[3], beta[3], dYZ, dZX, dXY;
  double tx, ty, tz;      // dt/dx, dt/dy, dt/dz
  double cX, cY, cZ, cs;  // dt/(mu0 eps0 d*), dt/eps0
};

template <class R>
struct CpmlArgsT {
  int enabled;
  int lowerZ, upperZ;
  int zb;                               // slid layer base of transverse histories
  const CpmlPointT<R>* prof[2][3];      // [magnetic/electric][axis]
  const int* ord[2][3];                 // ordinal of an active index, -1 if inactive
  R* hist[2][3][3];                     // [mag/ele][target comp][derivative]
  long long ostride[2][3][3];           // stride between ordinals
  int row[2][3][3];                     // row length of a layer
  double magScale[3], eleScale[3];
};

// Tile footprint of the last deposit: J is zero outside the listed cell
// tiles (those with an occupied tile within a deposit's reach), so
// clearing it touches only those.
struct CurrentFootprint {
  bool valid = false;
  const int* active = nullptr;
  int count = 0;
  int ntx = 0, nty = 0, ntz = 0, tx = 0, ty = 0, tz = 0;
};

// Cell-tile decomposition of the grid: ntx x nty x ntz tiles of
// tx x ty x tz cells; a tile owns the lattice nodes of its cells' lower
// corners, and the last tile of an axis also the closing node.
struct TileGrid {
  int ntx, nty, ntz, tx, ty, tz;
};

// Device scratch of the diagnostics.
struct DiagnosticScratch {
  DeviceArray<double> partial, out;
  DeviceArray<unsigned long long> acc, maxima;
};

// Moving-window field energy ledger (MovingWindowFieldEnergyLedger,
// moving_window_fields.h:39-58) on the device: every per-step measurement
// and accumulation is stream-ordered, nothing is read back until asked.
struct EnergyLedgerDev {
  bool enabled = false;
  bool stepFinished = true;
  // a charge-matched window measures the inserted plane after the inlet
  // completion that follows the shift
  bool insertPending = false;
  // compensated (sum, carry) pairs: kLedger* slots
  DeviceArray<double> acc;
  // per-step values: start power [0, 6), overwrite delta [6, 12), power
  // before the inlet overwrite [12, 18), power after / at the step end
  // [18, 24), plane energy before [24] and after [25] the overwrite,
  // removed plane energy [26], inserted plane energy [27]
  DeviceArray<double> work;
  unsigned long long overwriteEvents = 0, fluxSteps = 0, remapSteps = 0;
};

template <class R>
struct FieldStateT {
  aprl_grid_geometry g{};
  aprl_precision precision = APRL_FP64;
  CowanCoef cw{};
  std::size_t dims[9][3]{};
  std::size_t slack = 0, zb = 0;
  DeviceArray<R> comp[9];
  DeviceArray<R> scratch;

  // boundary
  bool hasCpml = false, lowerZ = true, upperZ = true;
  aprl_cpml_parameters cpml{};
  DeviceArray<CpmlPointT<R>> prof[2][3];
  DeviceArray<int> ord[2][3];
  DeviceArray<R> hist[12];
  std::size_t histSlots[12]{}, histLayers[12]{}, histRow[12]{};
  bool histSlides[12]{};

  // halos and moving window
  DeviceArray<R> haloLo[3], haloUp[3], zeroPlane[3];
  bool window = false;
  std::size_t step = 0;
  std::ptrdiff_t offset = 0;
  bool established = false;
  int phase = 0;  // 0 Idle, 1 Magnetic, 2 Electric, 3 Complete
  bool inletSet = false;
  DeviceArray<R> inlet[2], shiftPlanes[6];
  const R* loPlane[3] = {nullptr, nullptr, nullptr};  // lower E halo planes in use
  const R* inletPlane[2] = {nullptr, nullptr};         // inlet Ex, Ey planes in use
  // charge-matched inlet with the centered normal reference
  bool chargeMatch = false;
  DeviceArray<R> refEz;                      // reference Ez at the step's offset
  DeviceArray<unsigned long long> inletAcc;  // fixed-point inlet charge, (nx+1)(ny+1)
  DeviceArray<unsigned long long> inletWords;
  // the inlet reference on the device (aprl_window_set_reference): planes
  // not given by the caller come from it at the window's offset and step
  const aprl_inlet_reference* reference = nullptr;
  double referenceBeta = 0.0;

  // True while J is known to hold zeros on the device (set by clears and
  // shifts, dropped by any write of J), so repeated clears cost nothing.
  bool currentZero = true;
  CurrentFootprint footprint;

  hipStream_t stream = nullptr;
  Profiler profiler;
  DiagnosticScratch diag;
  EnergyLedgerDev ledger;

  std::size_t compSize(int a) const { return dims[a][0] * dims[a][1] * dims[a][2]; }
  std::size_t plane(int a) const { return dims[a][0] * dims[a][1]; }
  bool slides(int a) const { return a < 6; }
  R* base(int a) const { return comp[a].get() + (slides(a) ? zb * plane(a) : 0); }

  FieldPtrsT<R> ptrs() const;
  StepCoefsT<R> coefs() const;
  CpmlArgsT<R> cpmlArgs() const;
  std::size_t historyDeviceOffset(std::size_t h, std::size_t slot) const;
};

using CpmlPointDev = CpmlPointT<double>;
using FieldPtrs = FieldPtrsT<double>;
using StepCoefs = StepCoefsT<double>;
using CpmlArgs = CpmlArgsT<double>;
using FieldState = FieldStateT<double>;
using FieldState32 = FieldStateT<float>;

void validateGeometry(const aprl_grid_geometry& g, const CowanCoef& cw);
template <class R>
void allocateFields(FieldStateT<R>& s, std::size_t slack);
template <class R>
void installCpml(FieldStateT<R>& s, const aprl_cpml_parameters& p, std::size_t globalNz, std::size_t zOffset,
                 bool lowerZ, bool upperZ);

// Stream-ordered field operations.
template <class R>
void launchMagnetic(FieldStateT<R>& s, const FieldPtrsT<R>& f);
// mode 0: in place with history advance; mode 1: preview into `out` without
// touching histories (and, when sum != nullptr, sum = preview + old E); the
// window inlet overwrite applies in both modes. With a gate the launch does
// nothing unless the device word *gate is nonzero with its top bit clear
// when it runs (a commit enqueued ahead of the decision that enables it).
template <class R>
void launchElectric(FieldStateT<R>& s, const FieldPtrsT<R>& f, int mode, R* const out[3], const R* inletX,
                    const R* inletY, R* const sum[3] = nullptr, const unsigned* gate = nullptr);
// The MODE 1 preview of launchElectric (with sums) restricted to the nodes
// owned by the first count tiles of `tiles`, count = high word of *total
// (read on the device).
template <class R>
void launchElectricPreviewTiles(FieldStateT<R>& s, const FieldPtrsT<R>& f, R* const out[3], R* const sum[3],
                                const R* inletX, const R* inletY, const int* tiles, const unsigned long long* total,
                                const TileGrid& tg);
template <class R>
void launchClearCurrent(FieldStateT<R>& s);
template <class R>
void windowShift(FieldStateT<R>& s, const R* const planes[6]);

}  // namespace aprl_gpu
