// Oracle: the laboratory-plane input route of a moving-window run
// (TEST INFRASTRUCTURE ONLY; see the header of orc_fields.hpp).
//
// Restates, for one rank, in the reference's arithmetic order:
//   eb_bunch.cpp:134-289      SIBunchPreprocessor::projectLabPlaneEventsToSnapshot
//   eb_bunch.cpp:291-352      SIBunchPreprocessor::placeRelativeLabSnapshot
//   eb_bunch.cpp:354-482      SIBunchPreprocessor::boostLabSnapshotToBoxTimeZero
//   simulation.cpp:2126-2186  the frozen-injection frame origin (earliest crossing)
//   simulation.cpp:2228-2245  the frozen flags, held speeds and the upstream check
//   boostframe.cpp:74-164     gammaFromProperVelocity, longitudinalOnlyBetaZ,
//                             heldLongitudinalBetaZ, the light-front transforms
// Sums the reference keeps in long double are serial long-double sums here.
// Pinning: bitwise against the reference build's own SIBunchPreprocessor and
// BoostFrameTransform (oracle/_ref; tests/test_oracle_bunch.py), and by the
// closed forms the reference documents (the weighted longitudinal centroid
// lands on the reference position, a backward projection is rejected).
#pragma once

#include <cmath>
#include <iomanip>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <vector>

#include "orc_particles.hpp"

namespace orc {
namespace bunch {

// BoostFrameTransform with the reference's origin members and light-front
// momentum transforms (boostframe.cpp:42-60, 74-164, 172-240).
struct Frame {
  double gamma = 1.0, beta = 0.0, gammaBeta = 0.0, labT0 = 0.0, labZ0 = 0.0, boxZ0 = 0.0;
  static double gammaBetaFromGamma(double g) {
    if (!(g >= 1.0) || !std::isfinite(g)) throw std::invalid_argument("Lorentz gamma must be finite and >= 1");
    return std::sqrt(g - 1.0) * std::sqrt(g + 1.0);
  }
  void set(double g, double tLab, double zLab, double zBox) {
    gamma = g;
    beta = gammaBetaFromGamma(g) / g;
    gammaBeta = gammaBetaFromGamma(g);
    if (!std::isfinite(tLab) || !std::isfinite(zLab) || !std::isfinite(zBox))
      throw std::invalid_argument("Lorentz origins must be finite");
    labT0 = tLab;
    labZ0 = zLab;
    boxZ0 = zBox;
  }
  void labToBox(double tLab, double zLab, double& tBox, double& zBox) const {
    const double time = tLab - labT0, position = zLab - labZ0;
    tBox = std::fma(-gammaBeta / kC, position, gamma * time);
    zBox = boxZ0 + std::fma(-gammaBeta * kC, time, gamma * position);
  }
  double labZFromBoxZT(double zBox, double tBox) const {
    return labZ0 + std::fma(gammaBeta * kC, tBox, gamma * (zBox - boxZ0));
  }
  double labTFromBoxZT(double zBox, double tBox) const {
    return labT0 + std::fma(gammaBeta / kC, zBox - boxZ0, gamma * tBox);
  }
  static double positiveLightFront(const V3& u) {
    const double g = gammaFromU(u);
    const double mt2 = std::fma(u[0], u[0], std::fma(u[1], u[1], 1.0));
    if (!(mt2 > 0.0) || !std::isfinite(mt2))
      throw std::overflow_error("Cannot form a finite transverse mass for Lorentz transformation");
    if (u[2] >= 0.0) return g + u[2];
    return mt2 / (g - u[2]);
  }
  static V3 fromPositiveLightFront(const V3& u, double positive) {
    const double mt2 = std::fma(u[0], u[0], std::fma(u[1], u[1], 1.0));
    if (!(positive > 0.0) || !std::isfinite(positive))
      throw std::overflow_error("Lorentz transformation produced an invalid light-front momentum");
    V3 out = u;
    out[2] = 0.5 * (positive - mt2 / positive);
    if (!std::isfinite(out[2]))
      throw std::overflow_error("Lorentz transformation produced a non-finite longitudinal momentum");
    return out;
  }
  V3 labToBoxU(const V3& u) const { return fromPositiveLightFront(u, positiveLightFront(u) / (gamma + gammaBeta)); }
  V3 boxToLabU(const V3& u) const { return fromPositiveLightFront(u, positiveLightFront(u) * (gamma + gammaBeta)); }
  // heldLongitudinalBetaZ + longitudinalOnlyBetaZ, boostframe.cpp:90-107
  double heldBetaZ(const V3& uBox) const {
    const V3 lab = boxToLabU(uBox);
    V3 longitudinal = v3();
    longitudinal[2] = lab[2];
    const V3 t = labToBoxU(longitudinal);
    return t[2] / std::hypot(1.0, t[2]);
  }
};

// SILabPlaneProjectionReport (eb_bunch.h:36-47)
struct Projection {
  double inputPlaneLab = 0, referencePositionLab = 0, minimumForwardDistance = 0, maximumForwardDistance = 0,
         recommendedMinimumReferencePosition = 0, referenceTimeOffsetLab = 0, meanLongitudinalBeta = 0;
  unsigned long long particles = 0;
};

// projectLabPlaneEventsToSnapshot, eb_bunch.cpp:134-289: position[2] holds
// the arrival time offset on input and the relative lab z on output.
inline Projection projectLabPlaneEvents(std::vector<Particle>& P, double inputPlane, double reference,
                                        double offset, bool hold) {
  if (!std::isfinite(inputPlane) || !std::isfinite(reference) || !std::isfinite(offset))
    throw std::invalid_argument("Lab-plane event projection coordinates must be finite");
  if (P.empty()) throw std::invalid_argument("Cannot project an empty particle event set");
  long double weight = 0.0L, weightedBeta = 0.0L, weightedBetaTime = 0.0L;
  double latest = -std::numeric_limits<double>::infinity();
  int invalid = 0;
  for (const Particle& p : P) {
    const double timeOffset = p.position[2];
    const double gamma = std::sqrt(1.0 + norm2(p.u));
    const double betaZ = p.u[2] / gamma;
    if (!(p.weight > 0.0) || !std::isfinite(p.weight) || !(betaZ > 0.0) || !std::isfinite(betaZ) ||
        !std::isfinite(timeOffset))
      invalid = 1;
    weight += static_cast<long double>(p.weight);
    weightedBeta += static_cast<long double>(p.weight) * static_cast<long double>(betaZ);
    weightedBetaTime += static_cast<long double>(p.weight) * static_cast<long double>(betaZ) *
                        static_cast<long double>(timeOffset);
    latest = std::max(latest, timeOffset);
  }
  if (invalid)
    throw std::invalid_argument(
        "Lab-plane event input requires finite positive weights, finite arrival times, and positive longitudinal "
        "velocity");
  if (!(weight > 0.0L) || !(weightedBeta > 0.0L))
    throw std::invalid_argument("Lab-plane event input has an invalid global weight or longitudinal velocity sum");
  const long double meanBeta = weightedBeta / weight;
  const long double referenceTime =
      (static_cast<long double>(reference - inputPlane) / static_cast<long double>(kC) + weightedBetaTime / weight) /
      meanBeta;
  const double scale = std::max(1.0, std::max(std::abs(inputPlane), std::abs(reference)));
  const double tolerance = 128.0 * std::numeric_limits<double>::epsilon() * scale;
  if (referenceTime < static_cast<long double>(latest) - static_cast<long double>(tolerance / kC)) {
    const long double minimumCenter =
        static_cast<long double>(inputPlane) +
        static_cast<long double>(kC) * (meanBeta * static_cast<long double>(latest) - weightedBetaTime / weight);
    std::ostringstream m;
    m << std::setprecision(16)
      << "The reconstructed laboratory snapshot would require backward propagation of at least one HDF5 lab-plane "
         "event. Set beam.reference.initial_center_z to at least "
      << static_cast<double>(minimumCenter)
      << " m, move input_plane_z upstream, or verify the input arrival-time values and selected page.";
    throw std::invalid_argument(m.str());
  }
  double lo = std::numeric_limits<double>::infinity(), hi = -std::numeric_limits<double>::infinity();
  for (Particle& p : P) {
    const double timeOffset = p.position[2];
    const double gamma = std::sqrt(1.0 + norm2(p.u));
    const double betaZ = p.u[2] / gamma;
    const double deltaTime = static_cast<double>(referenceTime) - timeOffset;
    double distance = betaZ * kC * deltaTime;
    if (distance < 0.0 && distance >= -tolerance) distance = 0.0;
    lo = std::min(lo, distance);
    hi = std::max(hi, distance);
    if (!hold) {
      p.position[0] += p.u[0] / p.u[2] * distance;
      p.position[1] += p.u[1] / p.u[2] * distance;
    }
    p.position[2] = inputPlane + distance - reference + offset;
  }
  Projection r;
  r.inputPlaneLab = inputPlane;
  r.referencePositionLab = reference;
  r.minimumForwardDistance = std::max(0.0, lo);
  r.maximumForwardDistance = hi;
  r.recommendedMinimumReferencePosition = static_cast<double>(
      static_cast<long double>(inputPlane) +
      static_cast<long double>(kC) * (meanBeta * static_cast<long double>(latest) - weightedBetaTime / weight));
  r.referenceTimeOffsetLab = static_cast<double>(referenceTime);
  r.meanLongitudinalBeta = static_cast<double>(meanBeta);
  r.particles = P.size();
  return r;
}

// SIBunchPlacementReport (eb_bunch.h:23-34)
struct Placement {
  double relativeTailLab = 0, relativeHeadLab = 0, referencePositionLab = 0, headAfterLab = 0,
         actualHeadDistance = 0, recommendedMaximumReferencePosition = 0;
  unsigned long long particles = 0;
};

// placeRelativeLabSnapshot, eb_bunch.cpp:291-352.
inline Placement placeRelative(std::vector<Particle>& P, bool hasEntrance, double entrance, double reference,
                               double margin) {
  if ((hasEntrance && !std::isfinite(entrance)) || !std::isfinite(reference) || !std::isfinite(margin) ||
      margin < 0.0)
    throw std::invalid_argument("Relative bunch placement requires finite lab coordinates and a nonnegative margin");
  if (P.empty()) throw std::invalid_argument("Cannot place an empty particle bunch");
  double tail = std::numeric_limits<double>::infinity(), head = -std::numeric_limits<double>::infinity();
  for (const Particle& p : P) {
    if (!std::isfinite(p.position[2])) throw std::invalid_argument("Bunch position must be finite");
    tail = std::min(tail, p.position[2]);
    head = std::max(head, p.position[2]);
  }
  Placement r;
  r.particles = P.size();
  r.relativeTailLab = tail;
  r.relativeHeadLab = head;
  r.referencePositionLab = reference;
  r.headAfterLab = reference + head;
  if (hasEntrance) {
    r.actualHeadDistance = entrance - r.headAfterLab;
    r.recommendedMaximumReferencePosition = entrance - head - margin;
  } else {
    r.actualHeadDistance = std::numeric_limits<double>::infinity();
    r.recommendedMaximumReferencePosition = std::numeric_limits<double>::infinity();
  }
  for (Particle& p : P) p.position[2] += reference;
  return r;
}

// The frozen-injection frame origin, simulation.cpp:2126-2186: the earliest
// activation-plane crossing over the snapshot (lab z absolute), and from it
// the lab time origin (out[0]) and lab z origin (out[1]) of the frame.
inline void frozenFrameOrigin(const std::vector<Particle>& P, double gammaBoost, double inputPlane,
                              double activationPlane, double dz, bool hold, double out[2]) {
  double earliest = std::numeric_limits<double>::infinity();
  int invalidHeld = 0;
  Frame probe;
  if (hold) probe.set(gammaBoost, 0.0, activationPlane, 0.0);
  for (const Particle& p : P) {
    const double gamma = gammaFromU(p.u);
    const double betaZ = p.u[2] / gamma;
    const double snapshotZ = p.position[2];
    if (hold) {
      const V3 uBox = probe.labToBoxU(p.u);
      const double betaHeld = probe.heldBetaZ(uBox);
      const double labAdvance = probe.beta + betaHeld;
      if (!(labAdvance > 0.0) || !std::isfinite(labAdvance)) {
        invalidHeld = 1;
        continue;
      }
      const double inputTime = (inputPlane - snapshotZ) / (betaZ * kC);
      const double inputToActivation = inputPlane - activationPlane;
      const double crossingOriginTime =
          inputTime - inputToActivation * (1.0 + probe.beta * betaHeld) / (kC * labAdvance);
      const double twoCellMargin = 2.0 * dz / (probe.gamma * kC * labAdvance);
      earliest = std::min(earliest, crossingOriginTime - twoCellMargin);
    } else {
      earliest = std::min(earliest, (activationPlane - snapshotZ) / (betaZ * kC));
    }
  }
  if (invalidHeld)
    throw std::invalid_argument("Held-transverse pseudo-trajectory must advance in laboratory z for every particle");
  out[0] = hold ? earliest : earliest - 2.0 * dz / (gammaBoost * kC);
  out[1] = activationPlane;
}

// SIBunchBoostReport (eb_bunch.h:49-60)
struct BoostReport {
  double earliestLabEventTime = 0, latestLabEventTime = 0, maximumAbsoluteDriftTime = 0, minimumLabGamma = 0,
         maximumLabGamma = 0, maximumRelativeMomentumRoundTripError = 0, maximumRelativeGammaRoundTripError = 0;
};

// boostLabSnapshotToBoxTimeZero, eb_bunch.cpp:354-482 (the report's
// defaults from its constructor, eb_bunch.cpp:32-39: zero maxima, maximum
// gamma 1, +inf minimum gamma).
inline BoostReport boostToBoxTimeZero(std::vector<Particle>& P, const Frame& frame, double snapshotTime, bool hold,
                                      double inputPlane) {
  if (!std::isfinite(snapshotTime) || (hold && !std::isfinite(inputPlane)))
    throw std::invalid_argument("Bunch snapshot time and held-mode input plane must be finite");
  BoostReport r;
  r.minimumLabGamma = std::numeric_limits<double>::infinity();
  r.maximumLabGamma = 1.0;
  r.earliestLabEventTime = std::numeric_limits<double>::infinity();
  r.latestLabEventTime = -std::numeric_limits<double>::infinity();
  for (Particle& p : P) {
    const V3 uLab = p.u;
    const double gammaParticle = gammaFromU(p.u);
    V3 betaParticle = v3();
    for (int a = 0; a < 3; ++a) betaParticle[a] = (1.0 / gammaParticle) * p.u[a];
    double eventTime = 0.0, drift = 0.0, eventZBox = 0.0;
    V3 eventPosition = p.position;
    if (hold) {
      if (!(betaParticle[2] > 0.0))
        throw std::invalid_argument("Held-transverse input requires positive laboratory z velocity");
      const double inputTime = snapshotTime + (inputPlane - p.position[2]) / (kC * betaParticle[2]);
      double inputTimeBox = 0.0, inputZBox = 0.0;
      frame.labToBox(inputTime, inputPlane, inputTimeBox, inputZBox);
      const V3 uBox = frame.labToBoxU(uLab);
      const double betaHeld = frame.heldBetaZ(uBox);
      if (!(frame.beta + betaHeld > 0.0))
        throw std::invalid_argument("Held-transverse pseudo-trajectory does not advance in laboratory z");
      eventZBox = std::fma(-kC * betaHeld, inputTimeBox, inputZBox);
      eventTime = frame.labTFromBoxZT(eventZBox, 0.0);
      drift = eventTime - snapshotTime;
    } else {
      const double denominator = 1.0 - frame.beta * betaParticle[2];
      if (!(denominator > 0.0) || !std::isfinite(denominator))
        throw std::runtime_error("Particle does not intersect the requested boosted simultaneity plane");
      eventTime = (frame.labT0 + frame.beta * (p.position[2] - frame.labZ0) / kC -
                   frame.beta * betaParticle[2] * snapshotTime) /
                  denominator;
      drift = eventTime - snapshotTime;
      for (int a = 0; a < 3; ++a) eventPosition[a] += kC * drift * betaParticle[a];
      double eventTimeBox = 0.0;
      frame.labToBox(eventTime, eventPosition[2], eventTimeBox, eventZBox);
      const double timeTolerance =
          128.0 * std::numeric_limits<double>::epsilon() * std::max(1.0, std::abs(eventTime));
      if (std::abs(eventTimeBox) > timeTolerance)
        throw std::runtime_error("Free-drift bunch transform did not reach box time zero");
    }
    if (!hold) {
      p.position[0] = eventPosition[0];
      p.position[1] = eventPosition[1];
    }
    p.position[2] = eventZBox;
    p.u = frame.labToBoxU(uLab);
    const V3 back = frame.boxToLabU(p.u);
    V3 delta = back;
    for (int a = 0; a < 3; ++a) delta[a] -= uLab[a];
    const double momentumError = std::sqrt(norm2(delta)) / std::max(1.0, std::sqrt(norm2(uLab)));
    const double backGamma = gammaFromU(back);
    r.earliestLabEventTime = std::min(r.earliestLabEventTime, eventTime);
    r.latestLabEventTime = std::max(r.latestLabEventTime, eventTime);
    r.maximumAbsoluteDriftTime = std::max(r.maximumAbsoluteDriftTime, std::abs(drift));
    r.minimumLabGamma = std::min(r.minimumLabGamma, gammaParticle);
    r.maximumLabGamma = std::max(r.maximumLabGamma, gammaParticle);
    r.maximumRelativeMomentumRoundTripError = std::max(r.maximumRelativeMomentumRoundTripError, momentumError);
    r.maximumRelativeGammaRoundTripError =
        std::max(r.maximumRelativeGammaRoundTripError, std::abs(backGamma - gammaParticle) / gammaParticle);
  }
  if (P.empty()) {
    r.earliestLabEventTime = r.latestLabEventTime = 0.0;
    r.minimumLabGamma = std::numeric_limits<double>::infinity();
  }
  return r;
}

// The frozen flags of frozen lab-plane injection, simulation.cpp:2228-2245:
// every particle frozen, its held speed cached when the transverse motion
// is held, and none at or past the activation plane at box time zero.
inline void freeze(std::vector<Particle>& P, const Frame& frame, double activationPlane, bool hold) {
  int alreadyPast = 0;
  for (Particle& p : P) {
    p.frozen = true;
    if (hold) p.betaHeld = frame.heldBetaZ(p.u);
    if (frame.labZFromBoxZT(p.position[2], 0.0) >= activationPlane) alreadyPast = 1;
  }
  if (alreadyPast)
    throw std::runtime_error(
        "Frozen lab-plane injection requires every particle upstream of the input plane at box time zero");
}

}  // namespace bunch
}  // namespace orc
