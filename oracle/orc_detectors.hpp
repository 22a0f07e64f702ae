// Oracle restatement of the laboratory field-plane detector (TEST
// INFRASTRUCTURE ONLY; never linked into the product): the single-rank path
// of LabDetectorManager::sampleFieldPlanes (lab_detectors.cpp:2426-2487)
// with samplePlane (2836-2898, no external background: the simulation
// constructs the manager with a null background, simulation.cpp:3291-3293),
// sampleDetectorComponent / interpolationAxis (33-97), ownerRank
// (2811-2826), and the BoostFrameTransform members it calls
// (boostframe.cpp:36-72, 199-203, 242-246, 267-285).
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <vector>

#include "orc_fields.hpp"

namespace orc {

// BoostFrameTransform after setOriginsFromGamma (boostframe.cpp:42-72).
struct LabFrame {
  double gamma = 1.0, beta = 0.0, gammaBeta = 0.0, c0 = kC;
  double labT0 = 0.0, labZ0 = 0.0, boxZ0 = 0.0;

  static LabFrame fromGamma(double gamma, double labT0, double labZ0, double boxZ0) {
    if (!(gamma >= 1.0) || !std::isfinite(gamma)) throw std::invalid_argument("Lorentz gamma must be finite and >= 1");
    LabFrame f;
    const double gb = std::sqrt(gamma - 1.0) * std::sqrt(gamma + 1.0);
    f.gamma = gamma;
    f.beta = gb / gamma;
    f.gammaBeta = gb;
    f.labT0 = labT0;
    f.labZ0 = labZ0;
    f.boxZ0 = boxZ0;
    return f;
  }
  // boostframe.cpp:199-203
  double boxZFromLabZAndBoxT(double zLab, double tBox) const {
    return boxZ0 + (zLab - labZ0) / gamma - beta * c0 * tBox;
  }
  // boostframe.cpp:242-246
  double labTFromLabZBoxT(double zLab, double tBox) const { return labT0 + tBox / gamma + beta * (zLab - labZ0) / c0; }
  // boostframe.cpp:267-285
  void fieldsBoxToLab(const double eBox[3], const double bBox[3], double eLab[3], double bLab[3]) const {
    eLab[0] = std::fma(gammaBeta * c0, bBox[1], gamma * eBox[0]);
    eLab[1] = std::fma(-gammaBeta * c0, bBox[0], gamma * eBox[1]);
    eLab[2] = eBox[2];
    bLab[0] = std::fma(-gammaBeta / c0, eBox[1], gamma * bBox[0]);
    bLab[1] = std::fma(gammaBeta / c0, eBox[0], gamma * bBox[1]);
    bLab[2] = bBox[2];
  }
};

// lab_detectors.cpp:33-58
struct InterpolationAxis {
  std::size_t lower = 0;
  double fraction = 0.0;
};
inline InterpolationAxis interpolationAxis(double lattice, std::size_t size) {
  InterpolationAxis r;
  if (lattice <= 0.0) {
    r.lower = 0;
    r.fraction = 0.0;
  } else if (lattice >= static_cast<double>(size - 1)) {
    r.lower = size - 2;
    r.fraction = 1.0;
  } else {
    r.lower = static_cast<std::size_t>(std::floor(lattice));
    r.fraction = lattice - static_cast<double>(r.lower);
  }
  return r;
}

// lab_detectors.cpp:66-97, one rank (every plane local).
inline double sampleDetectorComponent(const Comp& c, double latticeX, double latticeY, const InterpolationAxis& z) {
  const InterpolationAxis x = interpolationAxis(latticeX, c.nx);
  const InterpolationAxis y = interpolationAxis(latticeY, c.ny);
  double value = 0.0;
  for (unsigned cz = 0; cz < 2; ++cz) {
    const double weightZ = cz ? z.fraction : 1.0 - z.fraction;
    if (weightZ == 0.0) continue;
    const std::size_t k = z.lower + cz;
    for (unsigned b = 0; b < 2; ++b)
      for (unsigned a = 0; a < 2; ++a) {
        const double weightX = a ? x.fraction : 1.0 - x.fraction;
        const double weightY = b ? y.fraction : 1.0 - y.fraction;
        value += weightX * weightY * weightZ * c(x.lower + a, y.lower + b, k);
      }
  }
  return value;
}

// ownerRank (lab_detectors.cpp:2811-2826) on one rank: whether the box
// plane lies on the grid, with the 64-epsilon tolerance.
inline bool planeOnGrid(const Grid& g, double boxZ, double originZ) {
  const double coordinate = (boxZ - originZ) / g.dz;
  const double tolerance = 64.0 * kEpsilon * std::max(1.0, std::abs(coordinate));
  return !(coordinate < -tolerance || coordinate > static_cast<double>(g.nz) + tolerance);
}

// samplePlane (lab_detectors.cpp:2836-2898) without a background: per cell
// centre (i, j), j-major, E then B in the lab frame.
inline void samplePlane(const Fields& f, const LabFrame& frame, double coordinate, const double origin[3],
                        double* records) {
  const Grid& g = f.g;
  const InterpolationAxis vertexZ = interpolationAxis(coordinate, g.nz + 1);
  const InterpolationAxis halfZ = interpolationAxis(coordinate - 0.5, g.nz);
  for (std::size_t j = 0; j < g.ny; ++j) {
    const double py = origin[1] + (static_cast<double>(j) + 0.5) * g.dy;
    const double y = (py - origin[1]) / g.dy;
    for (std::size_t i = 0; i < g.nx; ++i) {
      const double px = origin[0] + (static_cast<double>(i) + 0.5) * g.dx;
      const double x = (px - origin[0]) / g.dx;
      double eBox[3], bBox[3];
      eBox[0] = sampleDetectorComponent(f.f[0], x - 0.5, y, vertexZ);
      eBox[1] = sampleDetectorComponent(f.f[1], x, y - 0.5, vertexZ);
      eBox[2] = sampleDetectorComponent(f.f[2], x, y, halfZ);
      bBox[0] = sampleDetectorComponent(f.f[3], x, y - 0.5, halfZ);
      bBox[1] = sampleDetectorComponent(f.f[4], x - 0.5, y, halfZ);
      bBox[2] = sampleDetectorComponent(f.f[5], x - 0.5, y - 0.5, vertexZ);
      double* r = records + 6 * (j * g.nx + i);
      frame.fieldsBoxToLab(eBox, bBox, r, r + 3);
    }
  }
}

// One field plane's cadence and sample of sampleFieldPlanes
// (lab_detectors.cpp:2430-2484) on one rank: true when this call sampled.
struct LabPlaneState {
  double z = 0.0, rhythm = 0.0;
  bool sampled = false;
  double nextTimeLab = 0.0;

  bool sample(const Fields& f, const LabFrame& frame, double timeBox, const double origin[3], double* records,
              double& timeLab) {
    const double boxZ = frame.boxZFromLabZAndBoxT(z, timeBox);
    if (!planeOnGrid(f.g, boxZ, origin[2])) return false;
    timeLab = frame.labTFromLabZBoxT(z, timeBox);
    const double tolerance =
        64.0 * kEpsilon * std::max(rhythm, std::max(std::abs(timeLab), std::abs(nextTimeLab)));
    if (sampled && timeLab < nextTimeLab - tolerance) return false;
    const double coordinate = (boxZ - origin[2]) / f.g.dz;
    samplePlane(f, frame, coordinate, origin, records);
    if (!sampled) {
      sampled = true;
      nextTimeLab = timeLab + rhythm;
    } else {
      do { nextTimeLab += rhythm; } while (nextTimeLab <= timeLab + tolerance);
    }
    return true;
  }
};

}  // namespace orc
