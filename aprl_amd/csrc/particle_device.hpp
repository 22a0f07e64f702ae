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

// source_coupling_math::momentum, source_coupling_math.h:272-293.
__device__ void momentum(double u[3], const double e[3], const double mag[3], double q, double m, double h,
                         double g0, unsigned& err) {
  const double before[3] = {u[0], u[1], u[2]};
  const double ks = q * h / (2 * m * kC);
  const double kick[3] = {e[0] * ks, e[1] * ks, e[2] * ks};
  const double bs = q * h / (2 * m);
  const double b[3] = {mag[0] * bs, mag[1] * bs, mag[2] * bs};
  const double a[3] = {before[0] + kick[0], before[1] + kick[1], before[2] + kick[2]};
  const double axb[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  const double b2 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
  const double ka = kick[0] * a[0] + kick[1] * a[1] + kick[2] * a[2];
  const double kb = kick[0] * b[0] + kick[1] * b[1] + kick[2] * b[2];
  const double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  const double kcross = kick[0] * axb[0] + kick[1] * axb[1] + kick[2] * axb[2];
  const double first = (kick[0] * before[0] + kick[1] * before[1] + kick[2] * before[2]) / g0;
  const double kk = kick[0] * kick[0] + kick[1] * kick[1] + kick[2] * kick[2];
  double G = g0 + first + (kk + kcross / g0 - first * first) / g0;
  G = fmax(G, .5 * (g0 + 1.));
  const double eps = 2.220446049250313e-16;
  int iteration = 0;
  for (; iteration < 12; ++iteration) {
    const double den = G * G + b2, numerator = G * G * ka + G * kcross + ab * kb;
    const double fv = G * (G - g0) - numerator / den;
    const double derivative = 2 * G - g0 - ((2 * G * ka + kcross) * den - 2 * G * numerator) / (den * den);
    const double correction = fv / derivative;
    const double next = G - correction;
    if (!isfinite(next) || next < .5 * (g0 + 1.) - 1e-13) {
      err |= kErrRootNonphysical;
      return;
    }
    G = fmax(next, .5 * (g0 + 1.));
    if (fabs(correction) <= 8 * eps * fmax(1., G)) break;
  }
  if (iteration == 12) {
    err |= kErrRootNoConverge;
    return;
  }
  const double den = G * G + b2;
  for (int k = 0; k < 3; ++k) u[k] = 2 * (G * G * a[k] + ab * b[k] + G * axb[k]) / den - before[k];
}

// Push of source_coupling.cpp:90-104 for one particle over one step with the
// gathered fields usedE / usedB (fixed over the step): the substeps with
// prescribed magnets at the explicit midpoint, then the particle-boundary
// and domain checks of the deposition loop against the grid origin o.
// Positions are absolute box coordinates in metres.
__device__ void pushChord(const StepArgs& S, const double o[3], const double x0[3], const double u0[3], double q,
                          double m, const double usedE[3], const double usedB[3], double x[3], double u[3],
                          double& absGrid, unsigned& err) {
  if (S.window && x0[2] < o[2]) err |= kErrUpstream;
  for (int a = 0; a < 3; ++a) { x[a] = x0[a]; u[a] = u0[a]; }
  const double h = S.dt / S.substeps;
  absGrid = 0.0;
  for (int n = 0; n < S.substeps; ++n) {
    const double bx[3] = {x[0], x[1], x[2]}, bu[3] = {u[0], u[1], u[2]};
    const double g0 = gammaFromU(u[0], u[1], u[2]);
    double point[3] = {x[0], x[1], x[2]};
    const double ps = .5 * h * kC / g0;
    for (int a = 0; a < 3; ++a) point[a] += ps * u[a];
    double ee[3] = {0., 0., 0.}, bb[3] = {0., 0., 0.};
    prescribedBox(S, point, S.time + (n + .5) * h, ee, bb);
    double e[3] = {usedE[0], usedE[1], usedE[2]}, b[3] = {usedB[0], usedB[1], usedB[2]};
    for (int a = 0; a < 3; ++a) { e[a] += ee[a]; b[a] += bb[a]; }
    momentum(u, e, b, q, m, h, g0, err);
    const double g1 = gammaFromU(u[0], u[1], u[2]);
    double mean[3] = {bu[0], bu[1], bu[2]};
    for (int a = 0; a < 3; ++a) mean[a] += u[a];
    const double xs = h * kC / (g0 + g1);
    for (int a = 0; a < 3; ++a) x[a] += xs * mean[a];
    double w = 0.0;
    for (int a = 0; a < 3; ++a) w += q * usedE[a] * (x[a] - bx[a]);
    absGrid += fabs(w);
  }
  // Boundary events of the deposition loop (simulation.cpp:5229-5245).
  for (int a = 0; a < 3; ++a) {
    if (!isfinite(x0[a]) || !isfinite(x[a])) err |= kErrNonfinite;
    if (x0[a] < S.physLo[a] - S.cpmlTol[a] || x0[a] > S.physHi[a] + S.cpmlTol[a]) err |= kErrStartInCpml;
    if (S.cpmlCells[a] && (x[a] < S.physLo[a] || x[a] > S.physHi[a])) err |= kErrBoundaryEvent;
    if (x0[a] < S.outLo[a] - S.outTol[a] || x0[a] > S.outHi[a] + S.outTol[a]) err |= kErrStartOutside;
    if (x[a] < S.outLo[a] || x[a] > S.outHi[a]) err |= kErrBoundaryEvent;
    const double rel0 = x0[a] - o[a], rel1 = x[a] - o[a];
    if (rel0 < -S.posTol[a] || rel0 > S.extent[a] + S.posTol[a] || rel1 < -S.posTol[a] ||
        rel1 > S.extent[a] + S.posTol[a])
      err |= kErrEndpointOutside;
  }
  if (S.hasExit) {
    const BoostDev& F = S.frame;
    const double za = F.labZ0 + fma(F.gammaBeta * F.c0, S.time, F.gamma * (x0[2] - F.boxZ0)) - S.exitLabZ;
    const double zb = F.labZ0 + fma(F.gammaBeta * F.c0, S.time + S.dt, F.gamma * (x[2] - F.boxZ0)) - S.exitLabZ;
    if (za >= 0.) err |= kErrBeyondExit;
    if (zb >= 0. && zb > za) err |= kErrBoundaryEvent;
  }
}

}  // namespace aprl_gpu
