// Device-resident Yee field state and its host-side bookkeeping.
//
// E and B components carry `slack` spare z planes in front of the window so
// that the moving-window shift (moving_window_fields.cpp:482-498) becomes a
// base-pointer decrement plus one inserted plane; a full recentering copy
// happens once every `slack` steps. CPML histories with a transverse
// derivative axis get the same per-slot slack on their z layers.
#pragma once

#include "common.hpp"

namespace aprl_gpu {

struct CowanCoef {
  double alpha[3], beta[3], deltaXY, deltaYZ, deltaZX, ratio[3];
};
CowanCoef cowanForZDispersion(double dx, double dy, double dz);

struct CpmlPointDev { double ikm1, decay, drive; };

// Kernel-side view of the field arrays (base pointers already slid).
struct FieldPtrs {
  double* e[3];
  double* b[3];
  double* j[3];
  int n[9][3];          // (nx, ny, nz) per component
  const double* lo[3];  // lower E halo planes (null => PEC parity)
  const double* up[3];  // upper E halo planes
};

struct StepCoefs {
  double alpha[3], beta[3], dYZ, dZX, dXY;
  double tx, ty, tz;        // dt/dx, dt/dy, dt/dz
  double cX, cY, cZ, cs;    // dt/(mu0 eps0 d*), dt/eps0
};

struct CpmlArgs {
  int enabled;
  int lowerZ, upperZ;
  int zb;                                   // slid layer base of transverse histories
  const CpmlPointDev* prof[2][3];           // [magnetic/electric][axis]
  const int* ord[2][3];                     // ordinal of an active index, -1 if inactive
  double* hist[2][3][3];                    // [mag/ele][target comp][derivative]
  long long ostride[2][3][3];               // stride between ordinals
  int row[2][3][3];                         // row length of a layer
  double magScale[3], eleScale[3];
};

// Device scratch of the diagnostics.
struct DiagnosticScratch {
  DeviceArray<double> partial, out;
  DeviceArray<unsigned long long> acc, maxima;
};

struct FieldState {
  aprl_grid_geometry g{};
  aprl_precision precision = APRL_FP64;
  CowanCoef cw{};
  std::size_t dims[9][3]{};
  std::size_t slack = 0, zb = 0;
  DeviceArray<double> comp[9];
  DeviceArray<double> scratch;

  // boundary
  bool hasCpml = false, lowerZ = true, upperZ = true;
  aprl_cpml_parameters cpml{};
  DeviceArray<CpmlPointDev> prof[2][3];
  DeviceArray<int> ord[2][3];
  DeviceArray<double> hist[12];
  std::size_t histSlots[12]{}, histLayers[12]{}, histRow[12]{};
  bool histSlides[12]{};

  // halos and moving window
  DeviceArray<double> haloLo[3], haloUp[3], zeroPlane[3];
  bool window = false;
  std::size_t step = 0;
  std::ptrdiff_t offset = 0;
  bool established = false;
  int phase = 0;  // 0 Idle, 1 Magnetic, 2 Electric, 3 Complete
  bool inletSet = false;
  DeviceArray<double> inlet[2], shiftPlanes[6];
  const double* loPlane[3] = {nullptr, nullptr, nullptr};  // lower E halo planes in use
  const double* inletPlane[2] = {nullptr, nullptr};         // inlet Ex, Ey planes in use

  // True while J is known to hold zeros on the device (set by clears and
  // shifts, dropped by any write of J), so repeated clears cost nothing.
  bool currentZero = true;
  // Tile footprint of the last deposit: J is zero outside the cell tiles
  // with an occupied neighbour tile, so clearing it touches only those.
  struct CurrentFootprint {
    bool valid = false;
    const int* active = nullptr;  // tiles with an occupied neighbour tile
    int count = 0;
    int ntx = 0, nty = 0, ntz = 0, tx = 0, ty = 0, tz = 0;
  } footprint;

  hipStream_t stream = nullptr;
  Profiler profiler;
  DiagnosticScratch diag;

  std::size_t compSize(int a) const { return dims[a][0] * dims[a][1] * dims[a][2]; }
  std::size_t plane(int a) const { return dims[a][0] * dims[a][1]; }
  bool slides(int a) const { return a < 6; }
  double* base(int a) const { return comp[a].get() + (slides(a) ? zb * plane(a) : 0); }

  FieldPtrs ptrs() const;
  StepCoefs coefs() const;
  CpmlArgs cpmlArgs() const;
  std::size_t historyDeviceOffset(std::size_t h, std::size_t slot) const;
};

void validateGeometry(const aprl_grid_geometry& g, const CowanCoef& cw);
void allocateFields(FieldState& s, std::size_t slack);
void installCpml(FieldState& s, const aprl_cpml_parameters& p, std::size_t globalNz, std::size_t zOffset,
                 bool lowerZ, bool upperZ);

// Stream-ordered field operations.
void launchMagnetic(FieldState& s, const FieldPtrs& f);
// mode 0: in place with history advance; mode 1: preview into `out` without
// touching histories; the window inlet overwrite applies in both modes.
// mode 1 with sum != nullptr also writes sum = preview + old E.
void launchElectric(FieldState& s, const FieldPtrs& f, int mode, double* const out[3],
                    const double* inletX, const double* inletY, double* const sum[3] = nullptr);
void launchClearCurrent(FieldState& s);
void windowShift(FieldState& s, const double* const planes[6]);

}  // namespace aprl_gpu
