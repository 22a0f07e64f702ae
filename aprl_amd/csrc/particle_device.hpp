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
  double cpmlOuterLo[3], cpmlOuterHi[3];  // origin and origin + n h (CPML region)
  double outLo[3], outHi[3], outTol[3];   // origin and fma(n, h, origin) (open boundary)
  int cpmlCells[3], cpmlEnabled;
  double cpmlRate[3], cpmlThick[3], cpmlOrder;  // conductivity grading of the carriers' damping
  int hasExit, window;
  double exitLabZ;
  int hasActivation;
  double activationLabZ;
  int holdTransverse;  // held motion: z += c dt beta (1) or ballistic along u (0)
  double extent[3], posTol[3];  // validatePosition
};

// Per-step arguments of the push (time, magnets, boundary geometry); cp is
// the grid's CPML grading (null without a CPML).
StepArgs stepArgsFor(const CouplingState& C, const aprl_grid_geometry& g, const aprl_cpml_parameters* cp,
                     bool window, const double origin[3], double time, unsigned substeps);

__device__ __forceinline__ double gammaFromU(double ux, double uy, double uz) {
  const double scale = fmax(1.0, fmax(fabs(ux), fmax(fabs(uy), fabs(uz))));
  const double inv = 1.0 / scale;
  const double a = ux * inv, b = uy * inv, c = uz * inv;
  return scale * sqrt(inv * inv + a * a + b * b + c * c);
}

// SIMagneticElement::fieldsLab, eb_sources.cpp:346-446 (E is zero).
__device__ __forceinline__ void magnetLab(const MagnetDev& M, const double pos[3], double b[3]) {
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
__device__ __forceinline__ void prescribedBox(const StepArgs& S, const double posBox[3], double tBox, double e[3], double b[3]) {
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
__device__ __forceinline__ void momentum(double u[3], const double e[3], const double mag[3], double q, double m, double h,
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
__device__ __forceinline__ void pushChordAt(const StepArgs& S, double time, double dt, const double o[3], const double x0[3],
                            const double u0[3], double q, double m, const double usedE[3], const double usedB[3],
                            double x[3], double u[3], double& absGrid, unsigned& err) {
  if (S.window && x0[2] < o[2]) err |= kErrUpstream;
  for (int a = 0; a < 3; ++a) { x[a] = x0[a]; u[a] = u0[a]; }
  const double h = dt / S.substeps;
  absGrid = 0.0;
  for (int n = 0; n < S.substeps; ++n) {
    const double bx[3] = {x[0], x[1], x[2]}, bu[3] = {u[0], u[1], u[2]};
    const double g0 = gammaFromU(u[0], u[1], u[2]);
    double point[3] = {x[0], x[1], x[2]};
    const double ps = .5 * h * kC / g0;
    for (int a = 0; a < 3; ++a) point[a] += ps * u[a];
    double ee[3] = {0., 0., 0.}, bb[3] = {0., 0., 0.};
    prescribedBox(S, point, time + (n + .5) * h, ee, bb);
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
  // Start checks of the deposition loop's boundary queries
  // (particle_boundary.cpp:118-121, 325-330) and of validatePosition.
  for (int a = 0; a < 3; ++a) {
    if (!isfinite(x0[a]) || !isfinite(x[a])) err |= kErrNonfinite;
    if (x0[a] < S.physLo[a] - S.cpmlTol[a] || x0[a] > S.physHi[a] + S.cpmlTol[a]) err |= kErrStartInCpml;
    if (x0[a] < S.outLo[a] - S.outTol[a] || x0[a] > S.outHi[a] + S.outTol[a]) err |= kErrStartOutside;
    const double rel0 = x0[a] - o[a];
    if (rel0 < -S.posTol[a] || rel0 > S.extent[a] + S.posTol[a]) err |= kErrEndpointOutside;
  }
}

__device__ __forceinline__ void pushChord(const StepArgs& S, const double o[3], const double x0[3], const double u0[3],
                                          double q, double m, const double usedE[3], const double usedB[3],
                                          double x[3], double u[3], double& absGrid, unsigned& err) {
  pushChordAt(S, S.time, S.dt, o, x0, u0, q, m, usedE, usedB, x, u, absGrid, err);
}

// ---------------------------------------------------------------------------
// Particle-boundary events (the deposit loop of pushDepositAndMigrate,
// simulation.cpp:4310-4917, for one rank).

// Event code of one store entry for one trial (0: an active particle whose
// segment ends at its push end point).
constexpr int kEvtKindMask = 3;           // active segment end: 1 outer exit, 2 CPML entry, 3 window exit
constexpr int kEvtFaceShift = 2;          // face of that hit (3 bits)
constexpr int kEvtCarrier = 1 << 5;       // a carrier segment follows (always set for a carrier entry)
constexpr int kEvtCarrierExit = 1 << 6;   // the carrier segment ends at an outer face
constexpr int kEvtCarrierFaceShift = 7;   // face of that exit (3 bits)
constexpr int kEvtCarrierDrop = 1 << 10;  // the carrier's end weight is at or below the cutoff
constexpr int kEvtReleased = 1 << 11;     // a held particle was released during the step

struct BoundaryHit {
  double pos[3];
  double fraction;
  int face;
};

__device__ __forceinline__ double labZOf(const StepArgs& S, double z, double t) {
  const BoostDev& F = S.frame;
  return F.labZ0 + fma(F.gammaBeta * F.c0, t, F.gamma * (z - F.boxZ0));
}

// ChargeConservingCurrentDepositor::validatePosition, eb_deposition.cpp:228-247.
__device__ __forceinline__ void validateEnd(const StepArgs& S, const double o[3], const double x[3], unsigned& err) {
  for (int a = 0; a < 3; ++a) {
    const double rel = x[a] - o[a];
    if (!isfinite(x[a]) || rel < -S.posTol[a] || rel > S.extent[a] + S.posTol[a]) err |= kErrEndpointOutside;
  }
}

// ParticleCPMLRegion::firstEntry, particle_boundary.cpp:104-174 (start
// checks in pushChord).
__device__ __forceinline__ bool cpmlFirstEntry(const StepArgs& S, const double s[3], const double e[3], BoundaryHit& h,
                               unsigned& err) {
  if (!S.cpmlEnabled) return false;
  double best = INFINITY;
  int bestFace = 0;
  bool entered = false;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (!S.cpmlCells[a]) continue;
    double candidate = 0.0;
    int face = 0;
    bool axisEntered = false;
    if (e[a] < S.physLo[a]) {
      candidate = (S.physLo[a] - s[a]) / (e[a] - s[a]);
      face = 2 * a;
      axisEntered = true;
    } else if (e[a] > S.physHi[a]) {
      candidate = (S.physHi[a] - s[a]) / (e[a] - s[a]);
      face = 2 * a + 1;
      axisEntered = true;
    }
    if (axisEntered && candidate < best) {
      best = candidate;
      bestFace = face;
      entered = true;
    }
  }
  if (!entered) return false;
  const double ft = 64.0 * 2.220446049250313e-16;
  if (!isfinite(best) || best < -ft || best > 1.0 + ft) {
    err |= kErrCpmlLocate;
    return false;
  }
  h.fraction = fmax(0.0, fmin(1.0, best));
  h.face = bestFace;
  for (int a = 0; a < 3; ++a) {
    h.pos[a] = s[a] + h.fraction * (e[a] - s[a]);
    h.pos[a] = fmax(S.cpmlOuterLo[a], fmin(S.cpmlOuterHi[a], h.pos[a]));
  }
  // the normal coordinate sits on the face (static indices keep the hit in
  // registers)
#pragma unroll
  for (int a = 0; a < 3; ++a)
    if (a == bestFace / 2) h.pos[a] = (bestFace % 2) ? S.physHi[a] : S.physLo[a];
  return true;
}

// Simulation::firstMovingWindowEntry, simulation_moving_window.cpp:380-395
// (the lab exit plane only when one is set).
__device__ __forceinline__ bool windowFirstEntry(const StepArgs& S, const double s[3], const double e[3], double startFraction,
                                 BoundaryHit& h, unsigned& err) {
  bool present = cpmlFirstEntry(S, s, e, h, err);
  if (!S.hasExit) return present;
  const double a = labZOf(S, s[2], S.time + startFraction * S.dt) - S.exitLabZ;
  const double b = labZOf(S, e[2], S.time + S.dt) - S.exitLabZ;
  if (a >= 0.) err |= kErrBeyondExit;
  if (b >= 0. && b > a) {
    const double fraction = -a / (b - a);
    if (!present || fraction < h.fraction) {
      for (int d = 0; d < 3; ++d) h.pos[d] = s[d] + fraction * (e[d] - s[d]);
      h.fraction = fraction;
      h.face = 5;
      present = true;
    }
  }
  return present;
}

// ParticleOpenBoundary::firstExit, particle_boundary.cpp:303-373.
__device__ __forceinline__ bool outerFirstExit(const StepArgs& S, const double s[3], const double e[3], BoundaryHit& h,
                               unsigned& err) {
  double best = INFINITY;
  int bestFace = 0;
  bool escaped = false;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (!isfinite(s[a]) || !isfinite(e[a])) err |= kErrNonfinite;
    if (s[a] < S.outLo[a] - S.outTol[a] || s[a] > S.outHi[a] + S.outTol[a]) err |= kErrStartOutside;
    double candidate = 0.0;
    int face = 0;
    bool axisEscaped = false;
    if (e[a] < S.outLo[a]) {
      candidate = (S.outLo[a] - s[a]) / (e[a] - s[a]);
      face = 2 * a;
      axisEscaped = true;
    } else if (e[a] > S.outHi[a]) {
      candidate = (S.outHi[a] - s[a]) / (e[a] - s[a]);
      face = 2 * a + 1;
      axisEscaped = true;
    }
    if (axisEscaped && candidate < best) {
      best = candidate;
      bestFace = face;
      escaped = true;
    }
  }
  if (!escaped) return false;
  const double ft = 64.0 * 2.220446049250313e-16;
  if (best < -ft || best > 1.0 + ft || !isfinite(best)) {
    err |= kErrExitLocate;
    return false;
  }
  h.fraction = fmax(0.0, fmin(1.0, best));
  h.face = bestFace;
  for (int a = 0; a < 3; ++a) {
    h.pos[a] = s[a] + h.fraction * (e[a] - s[a]);
    h.pos[a] = fmax(S.outLo[a], fmin(S.outHi[a], h.pos[a]));
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
    if (a == bestFace / 2) h.pos[a] = (bestFace % 2) ? S.outHi[a] : S.outLo[a];
  return true;
}

// ParticleCPMLRegion::conductivityRate, particle_boundary.cpp:216-241.
__device__ __forceinline__ double conductivityRate(const StepArgs& S, const double p[3]) {
  double rate = 0.0;
  for (int a = 0; a < 3; ++a) {
    if (!S.cpmlCells[a]) continue;
    double depth = 0.0;
    if (p[a] < S.physLo[a]) depth = (S.physLo[a] - p[a]) / S.cpmlThick[a];
    else if (p[a] > S.physHi[a]) depth = (p[a] - S.physHi[a]) / S.cpmlThick[a];
    depth = fmax(0.0, fmin(1.0, depth));
    rate += S.cpmlRate[a] * pow(depth, S.cpmlOrder);
  }
  return rate;
}

// ParticleCPMLRegion::currentDampingFactor, particle_boundary.cpp:176-205:
// exp(-duration * mean rate) with the three-point Gauss-Legendre mean.
__device__ __forceinline__ double dampingFactor(const StepArgs& S, const double s[3], const double e[3], double duration) {
  if (duration == 0.0 || !S.cpmlEnabled) return 1.0;
  const double abscissa = sqrt(3.0 / 5.0);
  const double node[3] = {0.5 * (1.0 - abscissa), 0.5, 0.5 * (1.0 + abscissa)};
  const double weight[3] = {5.0 / 18.0, 4.0 / 9.0, 5.0 / 18.0};
  double average = 0.0;
  for (int k = 0; k < 3; ++k) {
    double p[3];
    for (int a = 0; a < 3; ++a) p[a] = s[a] + node[k] * (e[a] - s[a]);
    average += weight[k] * conductivityRate(S, p);
  }
  const double exponent = -duration * average;
  return exponent < log(2.2250738585072014e-308) ? 0.0 : exp(exponent);
}

// processCarrierSegment, simulation.cpp:4310-4407: the carrier moves from
// start (time fraction startFraction) to pos, is cut at its first outer
// exit, deposits with the midpoint damping and keeps the end damping.
// Returns the code bits of the carrier part; localEnd is its end as a
// fraction of start -> pos.
__device__ __forceinline__ int carrierSegment(const StepArgs& S, const double o[3], const double start[3], double startFraction,
                              const double pos[3], double weight, double& localEnd, double& wDeposit, double& wEnd,
                              unsigned& err) {
  BoundaryHit oh;
  const bool exits = outerFirstExit(S, start, pos, oh, err);
  localEnd = 1.0;
  double end[3] = {pos[0], pos[1], pos[2]};
  if (exits) {
    localEnd = oh.fraction;
    for (int a = 0; a < 3; ++a) end[a] = oh.pos[a];
  }
  const double globalEnd = startFraction + localEnd * (1.0 - startFraction);
  const double duration = (globalEnd - startFraction) * S.dt;
  double mid[3];
  for (int a = 0; a < 3; ++a) mid[a] = 0.5 * (start[a] + end[a]);
  wDeposit = weight * dampingFactor(S, start, mid, 0.5 * duration);
  wEnd = weight * dampingFactor(S, start, end, duration);
  validateEnd(S, o, start, err);
  validateEnd(S, o, end, err);
  int code = kEvtCarrier;
  if (exits) code |= kEvtCarrierExit | (oh.face << kEvtCarrierFaceShift);
  else if (wEnd <= 64.0 * 2.220446049250313e-16) code |= kEvtCarrierDrop;
  return code;
}

// End action of an active particle's segment x0 -> x1 starting at the
// step's time fraction sf (processActiveSegment, simulation.cpp:4787-4836
// without retirement planes or rank interfaces): a CPML entry or window exit
// at or before the first outer exit (within 64 eps) hands the rest of the
// segment to a new carrier of weight 1; an outer exit removes the particle.
// t1 / t2: segment fractions ending the unit-weight part and the carrier
// part.
__device__ __forceinline__ int activeSegmentEvents(const StepArgs& S, const double o[3], const double x0[3], const double x1[3],
                                   double sf, double& t1, double& t2, double& wDeposit, double& wEnd,
                                   unsigned& err) {
  BoundaryHit ch, oh;
  const bool entersCpml = S.window ? windowFirstEntry(S, x0, x1, sf, ch, err) : cpmlFirstEntry(S, x0, x1, ch, err);
  const bool exitsOuter = outerFirstExit(S, x0, x1, oh, err);
  const double tolerance = 64.0 * 2.220446049250313e-16;
  t1 = t2 = 1.0;
  wDeposit = wEnd = 0.0;
  if (entersCpml && (!exitsOuter || ch.fraction <= oh.fraction + tolerance)) {
    validateEnd(S, o, ch.pos, err);
    t1 = ch.fraction;
    const int kind = S.window && ch.face == 5 ? 3 : 2;
    double local = 1.0;
    const int code = kind | (ch.face << kEvtFaceShift) |
                     carrierSegment(S, o, ch.pos, sf + t1 * (1.0 - sf), x1, 1.0, local, wDeposit, wEnd, err);
    t2 = t1 + local * (1.0 - t1);
    return code;
  }
  if (exitsOuter) {
    validateEnd(S, o, oh.pos, err);
    t1 = t2 = oh.fraction;
    return 1 | (oh.face << kEvtFaceShift);
  }
  validateEnd(S, o, x1, err);
  return 0;
}

// ---------------------------------------------------------------------------
// Held particles (prepareParticle's frozen branch, simulation.cpp:
// 4951-5000): with hold_transverse_until_activation z += c dt beta_held and
// the transverse position stays; otherwise the held particle moves
// ballistically, x += (c dt / gamma) u (RelativisticBorisPusher::
// pushPosition, eb_particles.cpp:131-139). Released where its lab z reaches
// the activation plane, at the fraction f of the step, from the point the
// same motion reaches in f dt.

struct HeldMove {
  bool release;     // the lab z reaches the activation plane this step
  bool kink;        // released strictly inside the step
  double f;         // release time fraction
  double x1[3];     // held end (no release) or release point
  double remaining; // push time after the release
};

__device__ __forceinline__ void heldAdvance(const StepArgs& S, const double x0[3], const double u[3], double beta,
                                            double timeStep, double frac, double out[3]) {
  for (int a = 0; a < 3; ++a) out[a] = x0[a];
  if (S.holdTransverse) {
    // c dt beta, or c f dt beta at the release, in the reference's order
    out[2] += (frac < 0.0 ? kC * S.dt : kC * frac * S.dt) * beta;
  } else {
    const double s = kC * timeStep / gammaFromU(u[0], u[1], u[2]);
    for (int a = 0; a < 3; ++a) out[a] += s * u[a];
  }
}

__device__ __forceinline__ HeldMove heldMove(const StepArgs& S, const double x0[3], const double u[3], double beta) {
  HeldMove h;
  const double startLabZ = labZOf(S, x0[2], S.time);
  heldAdvance(S, x0, u, beta, S.dt, -1.0, h.x1);
  const double endLabZ = labZOf(S, h.x1[2], S.time + S.dt);
  h.release = S.hasActivation && endLabZ >= S.activationLabZ;
  h.f = 0.0;
  h.kink = false;
  h.remaining = 0.0;
  if (h.release) {
    h.f = startLabZ >= S.activationLabZ
              ? 0.0
              : fmax(0.0, fmin(1.0, (S.activationLabZ - startLabZ) / (endLabZ - startLabZ)));
    heldAdvance(S, x0, u, beta, h.f * S.dt, h.f, h.x1);
    const double activeStart = S.time + h.f * S.dt;
    h.kink = activeStart > S.time && activeStart < S.time + S.dt;
    h.remaining = (1.0 - h.f) * S.dt;
  }
  return h;
}

// One trial's move of a store entry (shared by both forms): a carrier's
// straight move; a held entry's held move, or its release (the frozen leg's
// checks, then the push of the rest of the step from the release point,
// processActiveSegment, simulation.cpp:4676-4746) and, upstream of the
// window, its admission (5121-5155); or the orbit push. Then the events of
// the active part (one call site, which keeps the kernel in registers).
// Returns the event code; t1 / t2 are fractions of the chord x0 -> x, or of
// kink -> x after a release kink.
__device__ __forceinline__ int entryStep(const StepArgs& S, const double o[3], int state, const double x0[3],
                                         const double u0[3], double q, double m, double weight, double beta,
                                         const double usedE[3], const double usedB[3], double x[3], double u[3],
                                         double& absGrid, double& t1, double& t2, double& wDeposit, double& wEnd,
                                         unsigned& err) {
  absGrid = 0.0;
  t1 = 0.0;
  t2 = 1.0;
  wDeposit = wEnd = 0.0;
  if (state == kStateCarrier) {
    const double scale = kC / gammaFromU(u0[0], u0[1], u0[2]);
    for (int a = 0; a < 3; ++a) {
      const double v = u0[a] * scale;
      x[a] = x0[a];
      x[a] += S.dt * v;
      u[a] = u0[a];
    }
    return carrierSegment(S, o, x0, 0.0, x, weight, t2, wDeposit, wEnd, err);
  }
  const bool held = state == kStateHeld;
  HeldMove h{};
  double xs[3] = {x0[0], x0[1], x0[2]};
  double t0 = S.time, dtp = S.dt;
  bool push = true;
  if (held) {
    h = heldMove(S, x0, u0, beta);
    for (int a = 0; a < 3; ++a) xs[a] = h.x1[a];
    push = h.release && h.remaining > 0.0;
    t0 = S.time + h.f * S.dt;
    dtp = h.remaining;
  }
  if (push) {
    pushChordAt(S, t0, dtp, o, xs, u0, q, m, usedE, usedB, x, u, absGrid, err);
  } else {
    for (int a = 0; a < 3; ++a) {
      x[a] = xs[a];
      u[a] = u0[a];
    }
  }
  // start of the event segment, its time fraction, and the map of its
  // fractions onto the deposited chord (t = ta + local tb)
  double sp[3] = {x0[0], x0[1], x0[2]};
  double sf = 0.0, ta = 0.0, tb = 1.0;
  int code = 0;
  if (held && h.release) {
    code = kEvtReleased;
    if (S.window && x0[2] < o[2]) err |= kErrHeldRelease;
    if (h.kink) {
      for (int a = 0; a < 3; ++a) sp[a] = h.x1[a];
      BoundaryHit b1, b2;
      if (cpmlFirstEntry(S, x0, sp, b1, err) || outerFirstExit(S, x0, sp, b2, err)) err |= kErrHeldFrozenLeg;
      validateEnd(S, o, x0, err);
      validateEnd(S, o, sp, err);
      sf = h.f;
    }
  } else if (held && S.window && x0[2] < o[2]) {
    // upstream: only the shaped support is deposited until the inlet
    t1 = t2 = 1.0;
    if (x[2] < o[2]) return 0;
    const double fraction = (o[2] - x0[2]) / (x[2] - x0[2]);
    if (!(fraction >= 0.0 && fraction <= 1.0) || !isfinite(fraction)) {
      err |= kErrAdmission;
      return 0;
    }
    for (int a = 0; a < 3; ++a) sp[a] = x0[a] + fraction * (x[a] - x0[a]);
    sp[2] = o[2];
    sf = fraction;
    ta = fraction;
    tb = 1.0 - fraction;
  }
  double l1, l2;
  code |= activeSegmentEvents(S, o, sp, x, sf, l1, l2, wDeposit, wEnd, err);
  t1 = ta + l1 * tb;
  t2 = ta + l2 * tb;
  return code;
}

// ---------------------------------------------------------------------------
// Entry classification of the per-step sort: removed entries and, in a
// moving window, the particle part of finishMovingWindowStep applied to the
// new origin (simulation_moving_window.cpp:353-367): carriers at or beyond
// the upper face are cropped, an active particle beyond it or a carrier
// below the origin is an error. Returns false for an entry keyed behind the
// live ones.
struct SortArgs {
  int window, nz;
  double lowZ, upZ, dz;    // window origin z, fma(nz, dz, origin z)
  unsigned* words;         // coupling error words: [0] errors, [2] keyed behind, [3] cropped
  unsigned* removed;       // the store's removed-entry count, cleared by the sort
};

__device__ __forceinline__ bool liveEntry(const SortArgs& K, int* state, int p, double z) {
  const int st = state[p];
  if (st < 0) {
    atomicAdd(K.words + 2, 1u);
    return false;
  }
  if (!K.window) return true;
  if (st == kStateCarrier) {
    if (z >= K.upZ) {
      state[p] = kStateCropped;
      atomicAdd(K.words + 2, 1u);
      atomicAdd(K.words + 3, 1u);
      return false;
    }
    if (z < K.lowZ) atomicOr(K.words, static_cast<unsigned>(kErrCarrierBehind));
    return true;
  }
  if (z < K.lowZ && st != kStateHeld) atomicOr(K.words, static_cast<unsigned>(kErrBehindInlet));
  // ownerRank (simulation.cpp:4068-4083) of one rank
  const double coordinate = (z - K.lowZ) / K.dz;
  const double tolerance = 64.0 * 2.220446049250313e-16 * fmax(1.0, fabs(coordinate));
  if (coordinate > static_cast<double>(K.nz) + tolerance) atomicOr(K.words, static_cast<unsigned>(kErrRemapOutside));
  return true;
}

// ---------------------------------------------------------------------------
// Commit of the converged trial's events and the counter reductions. Each
// block writes its per-slot charge sums and counts in entry order; the
// totals kernel adds the blocks in block order, so the sums are
// reproducible.

constexpr int kEventBlock = 256;

template <typename R>
__global__ void __launch_bounds__(kEventBlock) kCommitEvents(int n, const int* evt, int* state, R* w, const R* q,
                                                             R* wend, unsigned* removed, const unsigned* trialWord,
                                                             double* blockQ, unsigned* blockN) {
  __shared__ double sq[2][kEventBlock];
  __shared__ int ss[2][kEventBlock];
  if (!(*trialWord & kEvtAny)) return;
  const int t = static_cast<int>(threadIdx.x);
  const int p = blockIdx.x * kEventBlock + t;
  int s0 = -1, s1 = -1;
  double q0 = 0.0, q1 = 0.0;
  if (p < n) {
    const int e = evt[p];
    if (e) {
      const int kind = e & kEvtKindMask, face = (e >> kEvtFaceShift) & 7;
      const double charge = static_cast<double>(q[p]);
      bool dead = false;
      if (kind == 1) {
        s0 = 6 + face;
        q0 = charge;
        dead = true;
      } else if (kind == 2) {
        s0 = face;
        q0 = charge;
      } else if (kind == 3) {
        s0 = 18;
        q0 = charge;
      }
      const R startWeight = w[p];
      if (e & kEvtReleased) state[p] = kStateActive;
      if (e & kEvtCarrier) {
        const R we = wend[p];
        if (e & kEvtCarrierExit) {
          s1 = 12 + ((e >> kEvtCarrierFaceShift) & 7);
          q1 = charge * static_cast<double>(we);
          dead = true;
        } else if (e & kEvtCarrierDrop) {
          dead = true;
        } else {
          state[p] = kStateCarrier;
          w[p] = we;
        }
      }
      // keep the weight the entry started the step with (1 for a particle)
      const bool wasCarrier = (e & kEvtCarrier) && !kind && !(e & kEvtReleased);
      wend[p] = wasCarrier ? startWeight : R(1);
      if (dead) {
        state[p] = kStateRemoved;
        atomicAdd(removed, 1u);
      }
    }
  }
  sq[0][t] = q0;
  ss[0][t] = s0;
  sq[1][t] = q1;
  ss[1][t] = s1;
  __syncthreads();
  if (t < kCounterSlots) {
    double acc = 0.0;
    unsigned cnt = 0;
    const int k = t >= 12 && t < 18 ? 1 : 0;
    for (int i = 0; i < kEventBlock; ++i)
      if (ss[k][i] == t) {
        acc += sq[k][i];
        ++cnt;
      }
    blockQ[blockIdx.x * kCounterSlots + t] = acc;
    blockN[blockIdx.x * kCounterSlots + t] = cnt;
  }
}

// Cropped carriers (state kStateCropped) of the sorted tail [lo, n).
template <typename R>
__global__ void __launch_bounds__(kEventBlock) kCropCounters(int lo, int n, const int* state, const R* q, const R* w,
                                                             double* blockQ, unsigned* blockN) {
  __shared__ double sq[kEventBlock];
  __shared__ int ss[kEventBlock];
  const int t = static_cast<int>(threadIdx.x);
  const int p = lo + blockIdx.x * kEventBlock + t;
  const bool cropped = p < n && state[p] == kStateCropped;
  sq[t] = cropped ? static_cast<double>(q[p]) * static_cast<double>(w[p]) : 0.0;
  ss[t] = cropped ? 19 : -1;
  __syncthreads();
  if (t < kCounterSlots) {
    double acc = 0.0;
    unsigned cnt = 0;
    for (int i = 0; i < kEventBlock; ++i)
      if (ss[i] == t) {
        acc += sq[i];
        ++cnt;
      }
    blockQ[blockIdx.x * kCounterSlots + t] = acc;
    blockN[blockIdx.x * kCounterSlots + t] = cnt;
  }
}

}  // namespace aprl_gpu
