// Device math and step arguments shared by the FP64 and FP32-split particle
// paths: Lorentz factor, the prescribed lab magnets boosted to the box frame
// (evaluated in FP64 in both forms), and the per-step particle-boundary
// geometry.
#pragma once

#include "particles.hpp"

namespace aprl_gpu {

struct StepArgs {
  double time, dt;
  int substeps;
  BoostDev frame;
  int nMagnets;
  MagnetDev mag[kMaxMagnets];
  // particle boundary geometry (ParticleCPMLRegion / ParticleOpenBoundary)
  double physLo[3], physHi[3], cpmlTol[3];
  double outLo[3], outHi[3], outTol[3];
  int cpmlCells[3];
  int hasExit, window;
  double exitLabZ;
  double extent[3], posTol[3];  // validatePosition
};

// Per-step arguments of the push (time, magnets, boundary geometry).
StepArgs stepArgsFor(const CouplingState& C, const aprl_grid_geometry& g, bool window, const double origin[3],
                     double time, unsigned substeps);

__device__ __forceinline__ double gammaFromU(double ux, double uy, double uz) {
  const double scale = fmax(1.0, fmax(fabs(ux), fmax(fabs(uy), fabs(uz))));
  const double inv = 1.0 / scale;
  const double a = ux * inv, b = uy * inv, c = uz * inv;
  return scale * sqrt(inv * inv + a * a + b * b + c * c);
}

// SIMagneticElement::fieldsLab, eb_sources.cpp:346-446 (E is zero).
__device__ void magnetLab(const MagnetDev& M, const double pos[3], double b[3]) {
  b[0] = b[1] = b[2] = 0.0;
  const double rel0 = pos[0] - M.center[0], rel1 = pos[1] - M.center[1], z = pos[2] - M.center[2];
  if (M.type == 2) {
    if (z >= 0.0 && z <= M.length) {
      const double x = rel0 * M.cosA + rel1 * M.sinA;
      const double y = -rel0 * M.sinA + rel1 * M.cosA;
      const double bx = M.gradient * y, by = M.gradient * x;
      b[0] = bx * M.cosA - by * M.sinA;
      b[1] = bx * M.sinA + by * M.cosA;
    }
    return;
  }
  if (M.type == 1) {
    if (z >= 0.0 && z <= M.length) {
      b[0] = M.peak * M.cosA;
      b[1] = M.peak * M.sinA;
    }
    return;
  }
  const double twoPi = 2.0 * 3.141592653589793238462643383279502884;
  double ts = 0.0, ls = 0.0;
  if (z >= 0.0 && z <= M.length) {
    ts = sin(M.waveNumber * z);
    ls = cos(M.waveNumber * z);
    if (M.phase != 0.0) {
      ts = sin(M.waveNumber * z + M.phase);
      ls = cos(M.waveNumber * z + M.phase);
    }
    if (M.endRampPeriods > 0.0) {
      const double ramp = M.endRampPeriods * M.period;
      const double distance = fmin(z, M.length - z);
      if (distance < ramp) {
        const double ph = 0.5 * twoPi * distance / ramp;
        const double envelope = 0.5 * (1.0 - cos(ph));
        const double derivative = 0.25 * twoPi / ramp * sin(ph) * (z < ramp ? 1.0 : -1.0);
        ts = envelope * ts - derivative / M.waveNumber * ls;
        ls *= envelope;
      }
    }
  } else if (M.gaussianFringe) {
    const double fc = z < 0.0 ? z : z - M.length;
    const double distance = fabs(fc);
    if (distance >= M.fringeExtent) return;
    const double q = distance / M.fringeExtent;
    const double q2 = q * q, q3 = q2 * q, q4 = q3 * q, q5 = q4 * q;
    const double taper = 1.0 - 10.0 * q3 + 15.0 * q4 - 6.0 * q5;
    const double taperD = (-30.0 * q2 + 60.0 * q3 - 30.0 * q4) / M.fringeExtent;
    const double gaussian = exp(-0.5 * pow(M.waveNumber * distance, 2.0));
    ls = gaussian * taper;
    const double dd = gaussian * (taperD - M.waveNumber * M.waveNumber * distance * taper);
    ts = z < 0.0 ? dd / M.waveNumber : -dd / M.waveNumber;
  } else {
    return;
  }
  const double tc = rel0 * M.cosA + rel1 * M.sinA;
  const double tf = M.peak * cosh(M.waveNumber * tc) * ts;
  b[0] = tf * M.cosA;
  b[1] = tf * M.sinA;
  b[2] = M.peak * sinh(M.waveNumber * tc) * ls;
}

// SIFieldSourceSet::addPrescribedBox (eb_sources.cpp:571-582) with the
// boxToLab / fieldsLabToBox transforms (boostframe.cpp:188-196, 248-265).
__device__ void prescribedBox(const StepArgs& S, const double posBox[3], double tBox, double e[3], double b[3]) {
  if (S.nMagnets == 0) return;
  const BoostDev& F = S.frame;
  double posLab[3] = {posBox[0], posBox[1], posBox[2]};
  const double zShift = posBox[2] - F.boxZ0;
  posLab[2] = F.labZ0 + fma(F.gammaBeta * F.c0, tBox, F.gamma * zShift);
  double eL[3] = {0.0, 0.0, 0.0}, bL[3] = {0.0, 0.0, 0.0};
  for (int n = 0; n < S.nMagnets; ++n) {
    double cb[3];
    magnetLab(S.mag[n], posLab, cb);
    bL[0] += cb[0]; bL[1] += cb[1]; bL[2] += cb[2];
    eL[0] += 0.0; eL[1] += 0.0; eL[2] += 0.0;
  }
  double eb[3], bb[3];
  eb[0] = fma(-F.gammaBeta * F.c0, bL[1], F.gamma * eL[0]);
  eb[1] = fma(F.gammaBeta * F.c0, bL[0], F.gamma * eL[1]);
  eb[2] = eL[2];
  bb[0] = fma(F.gammaBeta / F.c0, eL[1], F.gamma * bL[0]);
  bb[1] = fma(-F.gammaBeta / F.c0, eL[0], F.gamma * bL[1]);
  bb[2] = bL[2];
  for (int a = 0; a < 3; ++a) { e[a] += eb[a]; b[a] += bb[a]; }
}

}  // namespace aprl_gpu
