// Oracle: particle side of APRL's conservative-source PIC step
// (TEST INFRASTRUCTURE ONLY; see the header of orc_fields.hpp).
//
// Restates, in the reference's association order:
//   boostframe.cpp:74-87, 188-196, 248-265   frame transforms
//   eb_sources.cpp:320-446, 539-582          prescribed lab magnets
//   longitudinal_shape.h:17-117, 172-187     shapes, weights, deposit
//   transverse_source_shape.h:8-10           x-binomial node weights
//   source_coupling_math.h:105-167, 235-293  gathers, clipping, momentum
//   source_coupling.cpp:31-51, 84-104, 105-159  coupling push/residual
//   simulation_source_coupling.cpp:16-65     Picard orchestration
//   simulation.cpp:4167-5605                 particle loop, active particles
// Pinning: bitwise against oracle/_ref for weights, gathers, momentum,
// deposit and whole coupled steps; plus charge continuity, gather/deposit
// adjointness and the Lapenta-Markidis identities (tests/test_oracle_*.py).
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <unordered_map>

#include "orc_fields.hpp"

namespace orc {

using V3 = std::array<double, 3>;
inline V3 v3(double a = 0.0) { return V3{{a, a, a}}; }
inline void pmv(V3& v, double s, const V3& x) { for (int i = 0; i < 3; ++i) v[i] += s * x[i]; }
inline void addTo(V3& v, const V3& x) { for (int i = 0; i < 3; ++i) v[i] += x[i]; }
inline void subFrom(V3& v, const V3& x) { for (int i = 0; i < 3; ++i) v[i] -= x[i]; }
inline void scaleBy(V3& v, double s) { for (int i = 0; i < 3; ++i) v[i] *= s; }
inline double norm2(const V3& v) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }
inline double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 cross(const V3& l, const V3& r) {
  V3 x = v3();
  x[0] = l[1] * r[2] - l[2] * r[1];
  x[1] = l[2] * r[0] - l[0] * r[2];
  x[2] = l[0] * r[1] - l[1] * r[0];
  return x;
}

constexpr double kPi = 3.141592653589793238462643383279502884;

// BoostFrameTransform::gammaFromProperVelocity, boostframe.cpp:74-87.
inline double gammaFromU(const V3& u) {
  const double scale = std::max(1.0, std::max(std::abs(u[0]), std::max(std::abs(u[1]), std::abs(u[2]))));
  const double inv = 1.0 / scale;
  const double ux = u[0] * inv, uy = u[1] * inv, uz = u[2] * inv;
  return scale * std::sqrt(inv * inv + ux * ux + uy * uy + uz * uz);
}

struct Boost {
  double gamma = 1.0, beta = 0.0, gammaBeta = 0.0, c0 = 1.0, labT0 = 0.0, labZ0 = 0.0, boxZ0 = 0.0;
  static double gammaBetaFromGamma(double g) {
    if (!(g >= 1.0) || !std::isfinite(g)) throw std::invalid_argument("Lorentz gamma must be finite and >= 1");
    return std::sqrt(g - 1.0) * std::sqrt(g + 1.0);
  }
  // setOriginsFromGamma, boostframe.cpp:42-60.
  void setOriginsFromGamma(double g, double c, double tLab, double zLab, double zBox) {
    const double b = gammaBetaFromGamma(g) / g;
    if (!std::isfinite(b) || std::abs(b) >= 1.0) throw std::invalid_argument("Lorentz beta must be finite and |beta| < 1");
    if (!(c > 0.0) || !std::isfinite(c)) throw std::invalid_argument("Transform light speed must be positive");
    gamma = g; beta = b; gammaBeta = g * b; c0 = c;
    gammaBeta = gammaBetaFromGamma(g);
    if (!std::isfinite(tLab) || !std::isfinite(zLab) || !std::isfinite(zBox))
      throw std::invalid_argument("Lorentz origins must be finite");
    labT0 = tLab; labZ0 = zLab; boxZ0 = zBox;
  }
  // boxToLab, boostframe.cpp:188-196.
  void boxToLab(double tBox, double zBox, double& tLab, double& zLab) const {
    const double zShift = zBox - boxZ0;
    zLab = labZ0 + std::fma(gammaBeta * c0, tBox, gamma * zShift);
    tLab = labT0 + std::fma(gammaBeta / c0, zShift, gamma * tBox);
  }
  double labZFromBoxZT(double zBox, double tBox) const {
    return labZ0 + std::fma(gammaBeta * c0, tBox, gamma * (zBox - boxZ0));
  }
  // fieldsLabToBox, boostframe.cpp:248-265.
  void fieldsLabToBox(const V3& eL, const V3& bL, V3& eB, V3& bB) const {
    eB[0] = std::fma(-gammaBeta * c0, bL[1], gamma * eL[0]);
    eB[1] = std::fma(gammaBeta * c0, bL[0], gamma * eL[1]);
    eB[2] = eL[2];
    bB[0] = std::fma(gammaBeta / c0, eL[1], gamma * bL[0]);
    bB[1] = std::fma(-gammaBeta / c0, eL[0], gamma * bL[1]);
    bB[2] = bL[2];
  }
};

enum MagnetType { kUndulator = 0, kDipole = 1, kQuadrupole = 2 };

// SIMagneticElement, eb_sources.cpp:320-446.
struct Magnet {
  int type = kUndulator;
  V3 center = v3();
  double length = 0, period = 0, peak = 0, angle = 0;
  bool gaussianFringe = true;
  double fringeCutoff = 1.0e-9, gradient = 0, endRampPeriods = 0, phase = 0;
  double waveNumber = 0, fringeExtent = 0, cosA = 1, sinA = 0;

  void prepare() {
    for (double c : center)
      if (!std::isfinite(c)) throw std::invalid_argument("Invalid SI magnetic element");
    if (!(length >= 0.0) || !std::isfinite(length) || !std::isfinite(peak) || !std::isfinite(angle))
      throw std::invalid_argument("Invalid SI magnetic element");
    if (type == kUndulator && (!(period > 0.0) || !std::isfinite(period)))
      throw std::invalid_argument("Undulator period must be positive");
    if (type == kUndulator && gaussianFringe &&
        (!(fringeCutoff > 0.0) || !(fringeCutoff < 1.0) || !std::isfinite(fringeCutoff)))
      throw std::invalid_argument("Undulator fringe_relative_cutoff must be in (0,1)");
    if (type == kDipole && !(length > 0.0)) throw std::invalid_argument("Dipole length must be positive");
    if (type == kQuadrupole && (!(length > 0.0) || !std::isfinite(gradient)))
      throw std::invalid_argument("Quadrupole length/gradient is invalid");
    if (!std::isfinite(endRampPeriods) || endRampPeriods < 0.0 ||
        (endRampPeriods > 0.0 && (type != kUndulator || gaussianFringe || 2.0 * endRampPeriods * period > length)))
      throw std::invalid_argument("Internal undulator end ramp requires non-Gaussian ends and fits inside length");
    if (!std::isfinite(phase) || (phase != 0.0 && (type != kUndulator || endRampPeriods <= 0.0)))
      throw std::invalid_argument("Undulator phase requires internal end ramps");
    const double twoPi = 2.0 * kPi;
    waveNumber = type == kUndulator ? twoPi / period : 0.0;
    fringeExtent = type == kUndulator && gaussianFringe ? std::sqrt(-2.0 * std::log(fringeCutoff)) / waveNumber : 0.0;
    cosA = std::cos(angle);
    sinA = std::sin(angle);
  }

  void fieldsLab(const V3& pos, V3& e, V3& b) const {
    e = v3();
    b = v3();
    V3 rel = pos;
    subFrom(rel, center);
    const double z = rel[2];
    if (type == kQuadrupole) {
      if (z >= 0.0 && z <= length) {
        const double x = rel[0] * cosA + rel[1] * sinA;
        const double y = -rel[0] * sinA + rel[1] * cosA;
        const double bx = gradient * y, by = gradient * x;
        b[0] = bx * cosA - by * sinA;
        b[1] = bx * sinA + by * cosA;
      }
      return;
    }
    if (type == kDipole) {
      if (z >= 0.0 && z <= length) {
        b[0] = peak * cosA;
        b[1] = peak * sinA;
      }
      return;
    }
    const double twoPi = 2.0 * kPi;
    double ts = 0.0, ls = 0.0;
    if (z >= 0.0 && z <= length) {
      ts = std::sin(waveNumber * z);
      ls = std::cos(waveNumber * z);
      if (phase != 0.0) {
        ts = std::sin(waveNumber * z + phase);
        ls = std::cos(waveNumber * z + phase);
      }
      if (endRampPeriods > 0.0) {
        const double ramp = endRampPeriods * period;
        const double distance = std::min(z, length - z);
        if (distance < ramp) {
          const double ph = 0.5 * twoPi * distance / ramp;
          const double envelope = 0.5 * (1.0 - std::cos(ph));
          const double derivative = 0.25 * twoPi / ramp * std::sin(ph) * (z < ramp ? 1.0 : -1.0);
          ts = envelope * ts - derivative / waveNumber * ls;
          ls *= envelope;
        }
      }
    } else if (gaussianFringe) {
      const double fc = z < 0.0 ? z : z - length;
      const double distance = std::abs(fc);
      if (distance >= fringeExtent) return;
      const double q = distance / fringeExtent;
      const double q2 = q * q, q3 = q2 * q, q4 = q3 * q, q5 = q4 * q;
      const double taper = 1.0 - 10.0 * q3 + 15.0 * q4 - 6.0 * q5;
      const double taperD = (-30.0 * q2 + 60.0 * q3 - 30.0 * q4) / fringeExtent;
      const double gaussian = std::exp(-0.5 * std::pow(waveNumber * distance, 2.0));
      ls = gaussian * taper;
      const double dd = gaussian * (taperD - waveNumber * waveNumber * distance * taper);
      ts = z < 0.0 ? dd / waveNumber : -dd / waveNumber;
    } else {
      return;
    }
    const double tc = rel[0] * cosA + rel[1] * sinA;
    const double tf = peak * std::cosh(waveNumber * tc) * ts;
    b[0] = tf * cosA;
    b[1] = tf * sinA;
    b[2] = peak * std::sinh(waveNumber * tc) * ls;
  }
};

// SIFieldSourceSet::addPrescribedBox with magnets only, eb_sources.cpp:539-582.
struct Sources {
  std::vector<Magnet> magnets;
  void addPrescribedBox(const V3& posBox, double tBox, const Boost& frame, V3& eBox, V3& bBox) const {
    if (magnets.empty()) return;
    V3 posLab = posBox;
    double tLab = 0.0;
    frame.boxToLab(tBox, posBox[2], tLab, posLab[2]);
    V3 eL = v3(), bL = v3();
    for (const Magnet& m : magnets) {
      V3 ce, cb;
      m.fieldsLab(posLab, ce, cb);
      addTo(eL, ce);
      addTo(bL, cb);
    }
    V3 ex = v3(), bx = v3();
    frame.fieldsLabToBox(eL, bL, ex, bx);
    addTo(eBox, ex);
    addTo(bBox, bx);
  }
};

enum Shape { kQuintic = 0, kNotch = 1, kBinomial3 = 2 };

// transverse_source_shape::nodeWeights, transverse_source_shape.h:8-10.
inline std::array<double, 4> nodeWeights(double f) {
  return {{0.25 * (1. - f), 0.5 - 0.25 * f, 0.25 + 0.25 * f, 0.25 * f}};
}

struct Basis { long long first; unsigned count; double v[12]; };

// longitudinal_shape::basis, longitudinal_shape.h:25-55.
inline Basis basis(double z, bool edge, int shape) {
  Basis r;
  std::fill(r.v, r.v + 12, 0.0);
  if (shape == kBinomial3) {
    const double cell = std::floor(z), f = z - cell;
    r.first = static_cast<long long>(cell) - 1;
    r.count = edge ? 3 : 4;
    const double c[3] = {.25, .5, .25};
    for (int k = -1; k <= 1; ++k) {
      const double w = c[k + 1];
      if (edge) r.v[k + 1] += w;
      else { r.v[k + 1] += w * (1. - f); r.v[k + 2] += w * f; }
    }
    return r;
  }
  if (shape == kNotch) {
    const double cell = std::floor(z), f = z - cell;
    r.first = static_cast<long long>(cell) - 3;
    r.count = edge ? 7 : 8;
    const double c[7] = {1. / 64., -6. / 64., 15. / 64., 44. / 64., 15. / 64., -6. / 64., 1. / 64.};
    for (int k = -3; k <= 3; ++k) {
      const double w = c[k + 3];
      if (edge) r.v[k + 3] += w;
      else { r.v[k + 3] += w * (1. - f); r.v[k + 4] += w * f; }
    }
    return r;
  }
  const unsigned degree = edge ? 4 : 5;
  const double cell = std::floor(z), f = z - cell;
  double n[6] = {1., 0., 0., 0., 0., 0.};
  for (unsigned j = 1; j <= degree; ++j) {
    double saved = 0.;
    for (unsigned k = 0; k < j; ++k) {
      const double t = n[k] / j;
      n[k] = saved + (k + 1 - f) * t;
      saved = (f + j - k - 1) * t;
    }
    n[j] = saved;
  }
  const double p = 5., a = p / 6., b = 3 * p / 40. + p * (p - 1) / 72.,
               d = 5 * p / 112. + p * (p - 1) / 80. + p * (p - 1) * (p - 2) / 1296.;
  const double c[7] = {-d / 64., b / 16. + 3 * d / 32., -a / 4. - b / 4. - 15 * d / 64., 1 + a / 2. + 3 * b / 8. + 5 * d / 16.,
                       -a / 4. - b / 4. - 15 * d / 64., b / 16. + 3 * d / 32., -d / 64.};
  r.first = static_cast<long long>(cell) - 5;
  r.count = degree + 7;
  for (unsigned j = 0; j <= degree; ++j)
    for (unsigned k = 0; k < 7; ++k) r.v[j + k] += n[j] * c[k];
  return r;
}

// binomialXZSegment, longitudinal_shape.h:61-94.
template <class Visit>
inline void binomialXZSegment(const Grid& grid, const double start[3], const double delta[3], double first,
                              double last, Visit& visit) {
  const double width = last - first;
  const double mid = (first + last) * 0.5;
  const double x = start[0] + mid * delta[0], y = start[1] + mid * delta[1], z = start[2] + mid * delta[2];
  long long i = static_cast<long long>(std::floor(x)), j = static_cast<long long>(std::floor(y));
  const long long k = static_cast<long long>(std::floor(z));
  if (i == static_cast<long long>(grid.nx)) --i;
  if (j == static_cast<long long>(grid.ny)) --j;
  const auto wx = nodeWeights(x - i);
  const auto wz = nodeWeights(z - k);
  const double wy[2] = {1. - (y - j), y - j};
  const double edge[3] = {0.25, 0.5, 0.25};
  const double slope[4] = {-0.25, -0.25, 0.25, 0.25};
  const double sy[2] = {-1., 1.};
  const double curvature = width * width / 12.;
  const double xy = curvature * delta[0] * delta[1];
  const double xz = curvature * delta[0] * delta[2];
  const double yz = curvature * delta[1] * delta[2];
  for (unsigned c = 0; c < 4; ++c) {
    for (unsigned b = 0; b < 2; ++b) {
      const double weight = width * (wy[b] * wz[c] + yz * sy[b] * slope[c]);
      for (unsigned a = 0; a < 3; ++a) visit(0, i - 1 + a, j + b, k - 1 + c, weight * edge[a]);
    }
    for (unsigned a = 0; a < 4; ++a) visit(1, i - 1 + a, j, k - 1 + c, width * (wx[a] * wz[c] + xz * slope[a] * slope[c]));
  }
  for (unsigned b = 0; b < 2; ++b)
    for (unsigned a = 0; a < 4; ++a) {
      const double weight = width * (wx[a] * wy[b] + xy * slope[a] * sy[b]);
      for (unsigned c = 0; c < 3; ++c) visit(2, i - 1 + a, j + b, k - 1 + c, weight * edge[c]);
    }
}

// longitudinal_shape::weights, longitudinal_shape.h:97-117.
template <class Visit>
inline unsigned weights(const Grid& g, const V3& origin, const V3& begin, const V3& end, Visit visit, int shape,
                        bool smoothX) {
  const double h[3] = {g.dx, g.dy, g.dz};
  double x[3], v[3], cuts[64] = {0., 1.};
  unsigned count = 2;
  for (unsigned a = 0; a < 3; ++a) {
    x[a] = (begin[a] - origin[a]) / h[a];
    v[a] = (end[a] - begin[a]) / h[a];
    if (v[a] != 0.)
      for (long long j = static_cast<long long>(std::floor(std::min(x[a], x[a] + v[a]))) + 1;
           j < std::max(x[a], x[a] + v[a]); ++j) {
        if (count == 64) throw std::runtime_error("Quintic orbit exceeds crossing budget");
        cuts[count++] = (j - x[a]) / v[a];
      }
  }
  std::sort(cuts, cuts + count);
  const double q[4] = {-.8611363115940525752, -.3399810435848562648, .3399810435848562648, .8611363115940525752};
  const double qw[4] = {.3478548451374538574, .6521451548625461426, .6521451548625461426, .3478548451374538574};
  unsigned segments = 0;
  for (unsigned s = 1; s < count; ++s) {
    const double width = cuts[s] - cuts[s - 1];
    if (width <= 0.) continue;
    ++segments;
    if (smoothX && shape == kBinomial3) {
      binomialXZSegment(g, x, v, cuts[s - 1], cuts[s], visit);
      continue;
    }
    const unsigned points = (shape == kNotch || shape == kBinomial3) ? 2 : 4;
    for (unsigned t = 0; t < points; ++t) {
      const double node = points == 2 ? (t ? 0.57735026918962576451 : -0.57735026918962576451) : q[t];
      const double weight = points == 2 ? 1. : qw[t];
      const double u = .5 * (cuts[s] + cuts[s - 1] + width * node), factor = .5 * width * weight;
      const double xx = x[0] + u * v[0], yy = x[1] + u * v[1], zz = x[2] + u * v[2];
      long long i = static_cast<long long>(std::floor(xx)), j = static_cast<long long>(std::floor(yy));
      if (i == static_cast<long long>(g.nx)) --i;
      if (j == static_cast<long long>(g.ny)) --j;
      const double fx = xx - i, fy = yy - j;
      const Basis nodes = basis(zz, false, shape), edge = basis(zz, true, shape);
      for (unsigned k = 0; k < nodes.count; ++k) {
        const double w = factor * nodes.v[k];
        if (w == 0.) continue;
        for (unsigned b = 0; b < 2; ++b) visit(0, i, j + b, nodes.first + k, w * (b ? fy : 1. - fy));
        for (unsigned a = 0; a < 2; ++a) visit(1, i + a, j, nodes.first + k, w * (a ? fx : 1. - fx));
      }
      for (unsigned k = 0; k < edge.count; ++k) {
        const double w = factor * edge.v[k];
        if (w == 0.) continue;
        for (unsigned b = 0; b < 2; ++b)
          for (unsigned a = 0; a < 2; ++a) visit(2, i + a, j + b, edge.first + k, w * (a ? fx : 1. - fx) * (b ? fy : 1. - fy));
      }
    }
  }
  return segments;
}

// clipOpenOrbit for one rank, source_coupling_math.h:234-250.
struct OpenOrbit { V3 begin, end; double first = 0., last = 1.; bool valid = true; };
inline OpenOrbit clipOpenOrbit(const Grid& g, const V3& origin, const V3& begin, const V3& end) {
  OpenOrbit r;
  r.begin = begin;
  r.end = end;
  const double ds[3] = {g.dx, g.dy, g.dz};
  const std::size_t size[3] = {g.nx, g.ny, g.nz};
  for (unsigned a = 0; a < 3; ++a) {
    const double x = (begin[a] - origin[a]) / ds[a], d = (end[a] - begin[a]) / ds[a];
    if (!std::isfinite(x) || !std::isfinite(d)) throw std::runtime_error("Nonfinite open coupling orbit");
    if (d == 0.) {
      if (x < 0. || x > size[a]) r.valid = false;
      continue;
    }
    {
      const double t = -x / d;
      if (d > 0.) r.first = std::max(r.first, t);
      else r.last = std::min(r.last, t);
    }
    {
      const double t = (double(size[a]) - x) / d;
      if (d > 0.) r.last = std::min(r.last, t);
      else r.first = std::max(r.first, t);
    }
  }
  r.valid = r.valid && r.last >= r.first;
  if (r.valid)
    for (unsigned a = 0; a < 3; ++a) {
      const double d = end[a] - begin[a];
      if (r.first != 0.) r.begin[a] = begin[a] + r.first * d;
      if (r.last != 1.) r.end[a] = begin[a] + r.last * d;
    }
  return r;
}

// NativeHalo::openMagneticValue on one rank: constant end continuation,
// source_coupling_math.h:113-119.
inline double openB(const Comp& v, long long i, long long j, long long k) {
  i = std::max(0LL, std::min(i, static_cast<long long>(v.nx) - 1));
  j = std::max(0LL, std::min(j, static_cast<long long>(v.ny) - 1));
  k = std::max(0LL, std::min(k, static_cast<long long>(v.nz) - 1));
  return v(i, j, k);
}

// magneticSample with openFaces and no bounds record, source_coupling_math.h:122-138.
inline V3 magneticSampleOpen(const Fields& f, const V3& origin, const V3& position) {
  V3 value = v3();
  const Grid& g = f.g;
  const double ds[3] = {g.dx, g.dy, g.dz};
  int cell[2][3];
  double fraction[2][3];
  for (unsigned half = 0; half < 2; ++half)
    for (unsigned d = 0; d < 3; ++d) {
      const double x = (position[d] - origin[d]) / ds[d] - .5 * half;
      cell[half][d] = static_cast<int>(std::floor(x));
      fraction[half][d] = x - cell[half][d];
    }
  for (unsigned axis = 0; axis < 3; ++axis) {
    const unsigned hx = axis == 0 ? 0 : 1, hy = axis == 1 ? 0 : 1, hz = axis == 2 ? 0 : 1;
    for (unsigned k = 0; k < 2; ++k)
      for (unsigned j = 0; j < 2; ++j)
        for (unsigned i = 0; i < 2; ++i) {
          const double x = i ? fraction[hx][0] : 1 - fraction[hx][0];
          const double y = j ? fraction[hy][1] : 1 - fraction[hy][1];
          const double z = k ? fraction[hz][2] : 1 - fraction[hz][2];
          const double b = openB(f.f[3 + axis], cell[hx][0] + int(i), cell[hy][1] + int(j), cell[hz][2] + int(k));
          value[axis] += x * y * z * b;
        }
  }
  return value;
}

// binomialXZMagneticSample, source_coupling_math.h:142-167.
inline V3 binomialXZMagneticSample(const Fields& f, const V3& origin, const V3& point) {
  const Grid& g = f.g;
  const double h[3] = {g.dx, g.dy, g.dz};
  V3 value = v3();
  for (unsigned a = 0; a < 3; ++a) {
    long long cell[3];
    double fraction[3];
    for (unsigned d = 0; d < 3; ++d) {
      const double pos = (point[d] - origin[d]) / h[d] - (a == d ? 0. : 0.5);
      cell[d] = static_cast<long long>(std::floor(pos));
      fraction[d] = pos - cell[d];
    }
    const auto wx = nodeWeights(fraction[0]);
    const auto wz = nodeWeights(fraction[2]);
    const double wy[2] = {1. - fraction[1], fraction[1]};
    for (unsigned k = 0; k < 4; ++k)
      for (unsigned j = 0; j < 2; ++j) {
        double row = 0.;
        for (unsigned i = 0; i < 4; ++i) row += wx[i] * openB(f.f[3 + a], cell[0] - 1 + i, cell[1] + j, cell[2] - 1 + k);
        value[a] += wz[k] * wy[j] * row;
      }
  }
  return value;
}

// source_coupling_math::momentum, source_coupling_math.h:272-293.
inline unsigned momentum(V3& u, const V3& e, const V3& magnetic, double q, double m, double h, double knownGamma) {
  const V3 before = u;
  const double g0 = knownGamma > 0. ? knownGamma : gammaFromU(before);
  V3 kick = e;
  scaleBy(kick, q * h / (2 * m * kC));
  V3 b = magnetic;
  scaleBy(b, q * h / (2 * m));
  V3 a = before;
  addTo(a, kick);
  const V3 axb = cross(a, b);
  const double b2 = dot(b, b), ka = dot(kick, a), kb = dot(kick, b), ab = dot(a, b), kcross = dot(kick, axb);
  const double first = dot(kick, before) / g0;
  double G = g0 + first + (dot(kick, kick) + kcross / g0 - first * first) / g0;
  G = std::max(G, .5 * (g0 + 1.));
  unsigned iteration = 0;
  for (; iteration < 12; ++iteration) {
    const double den = G * G + b2, numerator = G * G * ka + G * kcross + ab * kb;
    const double fv = G * (G - g0) - numerator / den;
    const double derivative = 2 * G - g0 - ((2 * G * ka + kcross) * den - 2 * G * numerator) / (den * den);
    const double correction = fv / derivative;
    const double next = G - correction;
    if (!std::isfinite(next) || next < .5 * (g0 + 1.) - 1e-13) throw std::runtime_error("Nonphysical local mean-gamma root");
    G = std::max(next, .5 * (g0 + 1.));
    if (std::abs(correction) <= 8 * kEpsilon * std::max(1., G)) break;
  }
  if (iteration == 12) throw std::runtime_error("Local mean-gamma root did not converge");
  const double den = G * G + b2;
  for (unsigned k = 0; k < 3; ++k) u[k] = 2 * (G * G * a[k] + ab * b[k] + G * axb[k]) / den - before[k];
  return iteration + 1;
}

inline long double chargeWork(double q, const V3& e, const V3& x, const V3& y) {
  long double s = 0;
  for (unsigned a = 0; a < 3; ++a) s += (long double)q * e[a] * ((long double)y[a] - x[a]);
  return s;
}

struct Particle {
  V3 position = v3(), u = v3();
  double charge = 0, mass = 0, weight = 1;
  std::uint64_t id = 0;
  // upstream lab-plane source state held until activation (eb_particles.h:21-22)
  bool frozen = false;
  double betaHeld = 0.0;
};

// The energy ledger's particle sums, Simulation (simulation.cpp:3591-3625),
// over particles_ (active and held), in long double as the reference sums
// them, with ParticleEnergyMoment::sample (particle_energy_audit.h:20-40):
// out = {kinetic, total, represented, rep * lab gamma, rep * box gamma,
// rep * lab z, lab kinetic, c p_z box, c p_z lab}.
inline V3 properVelocityBoxToLab(const Boost& frame, const V3& u) {
  // positiveLightFrontComponent / properVelocityFromPositiveLightFront,
  // boostframe.cpp:110-164
  const double gamma = gammaFromU(u);
  const double mt2 = std::fma(u[0], u[0], std::fma(u[1], u[1], 1.0));
  const double positive = (u[2] >= 0.0 ? gamma + u[2] : mt2 / (gamma - u[2])) * (frame.gamma + frame.gammaBeta);
  V3 lab = u;
  lab[2] = 0.5 * (positive - mt2 / positive);
  return lab;
}

inline void particleEnergySums(const std::vector<Particle>& particles, const Boost& frame, double timeBox,
                               long double out[9]) {
  for (int c = 0; c < 9; ++c) out[c] = 0.;
  const long double c = static_cast<long double>(kC);
  for (const Particle& p : particles) {
    const double gammaBox = gammaFromU(p.u);
    const long double mass = static_cast<long double>(p.mass);
    const long double rest = mass * c * c;
    const long double u2 = static_cast<long double>(p.u[0] * p.u[0] + p.u[1] * p.u[1] + p.u[2] * p.u[2]);
    out[0] += rest * u2 / static_cast<long double>(gammaBox + 1.0);
    out[1] += rest * static_cast<long double>(gammaBox);
    const long double represented = mass / static_cast<long double>(kMe);
    out[2] += represented;
    const V3 lab = properVelocityBoxToLab(frame, p.u);
    const double gammaLab = gammaFromU(lab);
    out[3] += represented * static_cast<long double>(gammaLab);
    out[4] += represented * static_cast<long double>(gammaBox);
    out[5] += represented * static_cast<long double>(frame.labZFromBoxZT(p.position[2], timeBox));
    // ParticleEnergyMoment::sample
    const long double restM = static_cast<long double>(p.mass) * kC * kC;
    long double ul2 = 0;
    for (unsigned a = 0; a < 3; ++a) ul2 += static_cast<long double>(lab[a]) * lab[a];
    const long double gammaL = std::sqrt(1 + ul2);
    out[6] += restM * ul2 / (gammaL + 1);
    out[7] += restM * p.u[2];
    out[8] += restM * lab[2];
  }
}

// ConservativeSourceCoupling (source_coupling.cpp) for one rank, with the
// longitudinal_shape::Current of depth 8 (production, compact=false).
class Coupling {
 public:
  Fields* fields = nullptr;
  int shape = kBinomial3;
  bool smoothX = true;
  V3 origin = v3();
  Comp old[3];
  Sources sources;
  Boost frame;
  bool replay = false;

  struct Slot {
    Particle first, last;
    V3 guess = v3(), usedE = v3(), usedB = v3(), nextE = v3(), nextB = v3();
    bool seen = false, initialized = false, replayed = false, nextReady = false;
    unsigned pushes = 0;
    double time = 0., dt = 0.;
    std::vector<Particle> states;
    std::vector<long double> gridWork;
  };
  std::unordered_map<std::uint64_t, std::size_t> index;
  std::vector<Slot> slots;

  bool wide() const { return true; }

  // begin, source_coupling.cpp:114.
  void begin(Fields& f, const V3& o, const std::vector<Particle>& p) {
    fields = &f;
    origin = o;
    for (int a = 0; a < 3; ++a) old[a] = f.f[a];
    index.clear();
    slots.clear();
    slots.reserve(p.size() + 32);
    for (const auto& q : p) {
      index.emplace(q.id, slots.size());
      slots.emplace_back();
    }
    replay = false;
  }
  Slot& slot(const Particle& p) {
    auto it = index.find(p.id);
    if (it != index.end()) return slots[it->second];
    index.emplace(p.id, slots.size());
    slots.emplace_back();
    return slots.back();
  }
  void beginTrial() { for (auto& q : slots) q.seen = false; }

  // Impl::electric (wide branch), source_coupling.cpp:74-82.
  V3 electric(const V3& a, const V3& b) {
    const Grid& g = fields->g;
    const OpenOrbit path = clipOpenOrbit(g, origin, a, b);
    if (!path.valid) throw std::out_of_range("Quintic active orbit outside domain");
    V3 result = v3();
    weights(g, origin, path.begin, path.end, [&](unsigned d, long long i, long long j, long long k, double w) {
      const Comp& v = fields->f[d];
      if (i < 0 || j < 0 || i >= (long long)v.nx || j >= (long long)v.ny) return;
      if (k < 0 || k >= (long long)v.nz) return;
      result[d] += .5 * w * (v(i, j, k) + old[d](i, j, k));
    }, shape, smoothX);
    scaleBy(result, path.last - path.first);
    return result;
  }
  // Impl::magnetic, source_coupling.cpp:83-94.
  V3 magnetic(const V3& a, const V3& b) {
    const Grid& g = fields->g;
    if (shape == kQuintic) {
      const OpenOrbit r = clipOpenOrbit(g, origin, a, b);
      if (!r.valid) throw std::out_of_range("Active coupling magnetic orbit has no in-domain portion");
      V3 middle = r.begin;
      addTo(middle, r.end);
      scaleBy(middle, .5);
      return magneticSampleOpen(*fields, origin, middle);
    }
    const OpenOrbit path = clipOpenOrbit(g, origin, a, b);
    if (!path.valid) throw std::out_of_range("Notch magnetic orbit outside domain");
    V3 middle = path.begin;
    addTo(middle, path.end);
    scaleBy(middle, .5);
    if (smoothX) return binomialXZMagneticSample(*fields, origin, middle);
    V3 result = v3();
    for (int k = -1; k <= 1; ++k) {
      V3 point = middle;
      point[2] += k * g.dz;
      const V3 value = magneticSampleOpen(*fields, origin, point);
      pmv(result, k == 0 ? .5 : .25, value);
    }
    return result;
  }

  // applyCurrentSource, source_coupling.cpp:127-132 (one rank).
  void applyCurrentSource(double sign) {
    const Grid& g = fields->g;
    for (unsigned a = 0; a < 3; ++a) {
      Comp& e = fields->f[a];
      const Comp& cur = fields->f[6 + a];
      for (std::size_t k = 0; k < e.nz; ++k)
        for (std::size_t y = 0; y < e.ny; ++y)
          for (std::size_t x = 0; x < e.nx; ++x) {
            const bool applied = a == 0 ? (y > 0 && y < g.ny && k != 0 && k != g.nz)
                                        : (a == 1 ? (x > 0 && x < g.nx && k != 0 && k != g.nz)
                                                  : (x > 0 && x < g.nx && y > 0 && y < g.ny));
            if (applied) {
              const std::size_t n = (k * e.ny + y) * e.nx + x;
              if (cur.v[n] != 0.) e.v[n] += sign * g.dt / kEps0 * cur.v[n];
            }
          }
    }
  }

  void swapElectric() { for (int a = 0; a < 3; ++a) std::swap(fields->f[a].v, old[a].v); }

  // push, source_coupling.cpp:133-147 (no observers or probes).
  void push(Particle& p, double time, double dt, unsigned steps) {
    Slot& q = slot(p);
    if (replay) {
      if (!q.seen || q.time != time || q.dt != dt || q.states.size() != steps + 1)
        throw std::logic_error("Missing converged particle replay");
      for (unsigned a = 0; a < 3; ++a)
        if (q.first.position[a] != p.position[a] || q.first.u[a] != p.u[a])
          throw std::logic_error("Release state changed during coupled replay");
      q.replayed = true;
      p = q.last;
      return;
    }
    if (!q.initialized) {
      q.guess = p.position;
      const double gamma = gammaFromU(p.u);
      pmv(q.guess, dt * kC / gamma, p.u);
      q.initialized = true;
    }
    q.first = p;
    q.time = time;
    q.dt = dt;
    q.seen = true;
    ++q.pushes;
    if (q.nextReady) {
      q.usedE = q.nextE;
      q.usedB = q.nextB;
    } else {
      q.usedE = electric(p.position, q.guess);
      q.usedB = magnetic(p.position, q.guess);
    }
    q.states.clear();
    q.gridWork.clear();
    q.states.push_back(p);
    const double h = dt / steps;
    for (unsigned n = 0; n < steps; ++n) {
      const Particle before = p;
      const double g0 = gammaFromU(p.u);
      V3 point = p.position;
      pmv(point, .5 * h * kC / g0, p.u);
      V3 ee = v3(), bb = v3();
      sources.addPrescribedBox(point, time + (n + .5) * h, frame, ee, bb);
      V3 e = q.usedE;
      addTo(e, ee);
      V3 b = q.usedB;
      addTo(b, bb);
      momentum(p.u, e, b, p.charge, p.mass, h, g0);
      const double g1 = gammaFromU(p.u);
      V3 mean = before.u;
      addTo(mean, p.u);
      pmv(p.position, h * kC / (g0 + g1), mean);
      q.gridWork.push_back(chargeWork(p.charge, q.usedE, before.position, p.position));
      q.states.push_back(p);
    }
    q.last = p;
    q.guess = p.position;
  }

  // localResidual, source_coupling.cpp:148-202 (serial branch).
  std::array<double, 4> localResidual() {
    double force = 0., defect = 0., scale = 0., calls = 0.;
    for (auto& s : slots) {
      if (!s.seen) continue;
      calls += 1;
      s.nextE = electric(s.first.position, s.last.position);
      s.nextB = magnetic(s.first.position, s.last.position);
      s.nextReady = true;
      V3 de = s.nextE, db = s.nextB;
      subFrom(de, s.usedE);
      subFrom(db, s.usedB);
      force = std::max(force, std::abs(s.first.charge / s.first.mass) * s.dt / kC *
                                  (std::sqrt(norm2(de)) + kC * std::sqrt(norm2(db))));
      defect += std::abs(chargeWork(s.first.charge, de, s.first.position, s.last.position));
      for (const auto& w : s.gridWork) scale += std::abs(w);
      scale += std::abs(chargeWork(s.first.charge, s.nextE, s.first.position, s.last.position));
    }
    return {{force, defect, scale, calls}};
  }

  // Current::deposit for one rank and depth 8, longitudinal_shape.h:175-187.
  unsigned deposit(const V3& begin, const V3& end, double charge) {
    const Grid& g = fields->g;
    double scale[3];
    for (unsigned a = 0; a < 3; ++a) scale[a] = charge * (end[a] - begin[a]) / (g.dx * g.dy * g.dz * g.dt);
    return weights(g, origin, begin, end, [&](unsigned a, long long i, long long j, long long k, double w) {
      if (scale[a] == 0.) return;
      Comp& v = fields->f[6 + a];
      if (i < 0 || i >= (long long)v.nx) {
        if (smoothX && shape == kBinomial3) return;
        throw std::out_of_range("Transverse current support outside x grid");
      }
      if (j < 0 || j >= (long long)v.ny) throw std::out_of_range("Transverse current support outside y grid");
      const double value = scale[a] * w;
      if (k >= 0 && k < (long long)v.nz) v(i, j, k) += value;
    }, shape, smoothX);
  }

  // ChargeConservingCurrentDepositor::validatePosition, eb_deposition.cpp:228-247.
  void validatePosition(const V3& p) const {
    const Grid& g = fields->g;
    const double extent[3] = {static_cast<double>(g.nx) * g.dx, static_cast<double>(g.ny) * g.dy,
                              static_cast<double>(g.nz) * g.dz};
    for (unsigned a = 0; a < 3; ++a) {
      const double rel = p[a] - origin[a];
      const double tol = 64.0 * kEpsilon * std::max(1.0, std::max(std::abs(origin[a]), extent[a]));
      if (!std::isfinite(p[a]) || rel < -tol || rel > extent[a] + tol)
        throw std::out_of_range("Particle trajectory endpoint is outside the E/B domain");
    }
  }
};

// Particle-boundary geometry of one step: the transverse CPML inner faces
// (ParticleCPMLRegion, particle_boundary.cpp:56-241), the outer faces
// (ParticleOpenBoundary::firstExit, 274-373) and the moving-window lab exit
// plane (Simulation::firstMovingWindowEntry, simulation_moving_window.cpp:380-395).
// Faces are numbered x-, x+, y-, y+, z-, z+ (ParticleBoundaryFace).
struct Hit {
  V3 position = v3();
  double fraction = 1.0;
  int face = 0;
};

struct ParticleBoundary {
  double physLower[3], physUpper[3], outerLower[3], outerUpper[3], tol[3];
  double exitLower[3], exitUpper[3], exitTol[3];  // ParticleOpenBoundary (fma upper face)
  double maxRate[3] = {0, 0, 0}, order = 3.0;
  std::size_t cells[3];
  double exitLabZ = std::numeric_limits<double>::infinity();
  bool window = false;

  // ParticleCPMLRegion constructor (particle_boundary.cpp:56-84) and
  // ParticleOpenBoundary constructor (254-287); params carries the CPML
  // grading (polynomial order, target reflection) and the particle cells.
  void set(const Grid& g, const V3& origin, const CpmlParams& params) {
    const double sp[3] = {g.dx, g.dy, g.dz};
    const std::size_t n[3] = {g.nx, g.ny, g.nz};
    order = params.order;
    for (unsigned a = 0; a < 3; ++a) {
      cells[a] = params.cells[a];
      if (2 * cells[a] >= n[a] && cells[a] > 0)
        throw std::invalid_argument("Particle CPML region must leave a physical interior");
      outerLower[a] = origin[a];
      outerUpper[a] = origin[a] + static_cast<double>(n[a]) * sp[a];
      physLower[a] = outerLower[a] + static_cast<double>(cells[a]) * sp[a];
      physUpper[a] = outerUpper[a] - static_cast<double>(cells[a]) * sp[a];
      tol[a] = 64.0 * kEpsilon * std::max(1.0, std::max(std::abs(outerLower[a]), std::abs(outerUpper[a])));
      maxRate[a] = params.maxRate(a, g);
      exitLower[a] = origin[a];
      exitUpper[a] = std::fma(static_cast<double>(n[a]), sp[a], origin[a]);
      exitTol[a] = 64.0 * kEpsilon * std::max(1.0, std::max(std::abs(exitLower[a]), std::abs(exitUpper[a])));
    }
  }
  bool cpmlEnabled() const { return cells[0] > 0 || cells[1] > 0 || cells[2] > 0; }

  // ParticleCPMLRegion::firstEntry, particle_boundary.cpp:104-174.
  bool firstEntry(const V3& s, const V3& e, Hit& hit) const {
    if (!cpmlEnabled()) return false;
    double best = std::numeric_limits<double>::infinity();
    int bestFace = 0;
    bool entered = false;
    for (unsigned a = 0; a < 3; ++a) {
      if (!std::isfinite(s[a]) || !std::isfinite(e[a]))
        throw std::runtime_error("Particle CPML trajectory contains a non-finite coordinate");
      if (s[a] < physLower[a] - tol[a] || s[a] > physUpper[a] + tol[a])
        throw std::out_of_range("Active particle starts inside the non-physical CPML region");
      if (cells[a] == 0) continue;
      double candidate = 0.0;
      int face = 0;
      bool axisEntered = false;
      if (e[a] < physLower[a]) {
        candidate = (physLower[a] - s[a]) / (e[a] - s[a]);
        face = static_cast<int>(2 * a);
        axisEntered = true;
      } else if (e[a] > physUpper[a]) {
        candidate = (physUpper[a] - s[a]) / (e[a] - s[a]);
        face = static_cast<int>(2 * a + 1);
        axisEntered = true;
      }
      if (axisEntered && candidate < best) {
        best = candidate;
        bestFace = face;
        entered = true;
      }
    }
    if (!entered) return false;
    const double ft = 64.0 * kEpsilon;
    if (!std::isfinite(best) || best < -ft || best > 1.0 + ft)
      throw std::runtime_error("Cannot locate the first particle entry into CPML");
    hit.fraction = std::max(0.0, std::min(1.0, best));
    hit.face = bestFace;
    for (unsigned a = 0; a < 3; ++a) {
      hit.position[a] = s[a] + hit.fraction * (e[a] - s[a]);
      hit.position[a] = std::max(outerLower[a], std::min(outerUpper[a], hit.position[a]));
    }
    const unsigned normal = static_cast<unsigned>(hit.face / 2);
    hit.position[normal] = (hit.face % 2) != 0 ? physUpper[normal] : physLower[normal];
    return true;
  }

  // Simulation::firstMovingWindowEntry, simulation_moving_window.cpp:380-395.
  bool windowEntry(const V3& s, const V3& e, double tStart, double tEnd, const Boost& frame, Hit& hit) const {
    bool present = firstEntry(s, e, hit);
    if (!window || !std::isfinite(exitLabZ)) return present;
    const double a = frame.labZFromBoxZT(s[2], tStart) - exitLabZ;
    const double b = frame.labZFromBoxZT(e[2], tEnd) - exitLabZ;
    if (a >= 0.) throw std::runtime_error("Active particle started beyond trusted moving-window exit");
    if (b >= 0. && b > a) {
      const double fraction = -a / (b - a);
      if (!present || fraction < hit.fraction) {
        for (unsigned d = 0; d < 3; ++d) hit.position[d] = s[d] + fraction * (e[d] - s[d]);
        hit.fraction = fraction;
        hit.face = 5;
        present = true;
      }
    }
    return present;
  }

  // ParticleOpenBoundary::firstExit, particle_boundary.cpp:303-373.
  bool firstExit(const V3& s, const V3& e, Hit& hit) const {
    double best = std::numeric_limits<double>::infinity();
    int bestFace = 0;
    bool escaped = false;
    for (unsigned a = 0; a < 3; ++a) {
      if (!std::isfinite(s[a]) || !std::isfinite(e[a]))
        throw std::runtime_error("Particle trajectory contains a non-finite coordinate");
      if (s[a] < exitLower[a] - exitTol[a] || s[a] > exitUpper[a] + exitTol[a])
        throw std::out_of_range("Particle trajectory starts outside the global E/B domain");
      double candidate = 0.0;
      int face = 0;
      bool axisEscaped = false;
      if (e[a] < exitLower[a]) {
        candidate = (exitLower[a] - s[a]) / (e[a] - s[a]);
        face = static_cast<int>(2 * a);
        axisEscaped = true;
      } else if (e[a] > exitUpper[a]) {
        candidate = (exitUpper[a] - s[a]) / (e[a] - s[a]);
        face = static_cast<int>(2 * a + 1);
        axisEscaped = true;
      }
      if (axisEscaped && candidate < best) {
        best = candidate;
        bestFace = face;
        escaped = true;
      }
    }
    if (!escaped) return false;
    const double ft = 64.0 * kEpsilon;
    if (best < -ft || best > 1.0 + ft || !std::isfinite(best))
      throw std::runtime_error("Cannot locate the first particle-domain exit");
    hit.fraction = std::max(0.0, std::min(1.0, best));
    hit.face = bestFace;
    for (unsigned a = 0; a < 3; ++a) {
      hit.position[a] = s[a] + hit.fraction * (e[a] - s[a]);
      hit.position[a] = std::max(exitLower[a], std::min(exitUpper[a], hit.position[a]));
    }
    const unsigned normal = static_cast<unsigned>(hit.face / 2);
    hit.position[normal] = (hit.face % 2) != 0 ? exitUpper[normal] : exitLower[normal];
    return true;
  }

  // ParticleCPMLRegion::conductivityRate, particle_boundary.cpp:216-241.
  double conductivityRate(const V3& p, const Grid& g) const {
    const double sp[3] = {g.dx, g.dy, g.dz};
    double rate = 0.0;
    for (unsigned a = 0; a < 3; ++a) {
      if (cells[a] == 0) continue;
      const double thickness = static_cast<double>(cells[a]) * sp[a];
      double depth = 0.0;
      if (p[a] < physLower[a]) depth = (physLower[a] - p[a]) / thickness;
      else if (p[a] > physUpper[a]) depth = (p[a] - physUpper[a]) / thickness;
      depth = std::max(0.0, std::min(1.0, depth));
      rate += maxRate[a] * std::pow(depth, order);
    }
    return rate;
  }

  // ParticleCPMLRegion::currentDampingFactor, particle_boundary.cpp:176-205.
  double damping(const V3& s, const V3& e, double duration, const Grid& g) const {
    if (!(duration >= 0.0) || !std::isfinite(duration))
      throw std::invalid_argument("CPML carrier segment duration must be finite and nonnegative");
    if (duration == 0.0 || !cpmlEnabled()) return 1.0;
    const double abscissa = std::sqrt(3.0 / 5.0);
    const double node[3] = {0.5 * (1.0 - abscissa), 0.5, 0.5 * (1.0 + abscissa)};
    const double weight[3] = {5.0 / 18.0, 4.0 / 9.0, 5.0 / 18.0};
    double average = 0.0;
    for (unsigned k = 0; k < 3; ++k) {
      V3 p = v3();
      for (unsigned a = 0; a < 3; ++a) p[a] = s[a] + node[k] * (e[a] - s[a]);
      average += weight[k] * conductivityRate(p, g);
    }
    const double exponent = -duration * average;
    return exponent < std::log(std::numeric_limits<double>::min()) ? 0.0 : std::exp(exponent);
  }
};

// ParticleCPMLCarrier (particle_boundary.h:76-84): a particle that entered
// the CPML (or passed the moving-window exit plane) moves on straight and
// deposits damped current. id and u are bookkeeping only: the particle the
// carrier came from and its proper velocity, from which velocity = u c/gamma.
struct Carrier {
  V3 position = v3(), velocity = v3(), u = v3();
  double charge = 0, weight = 1;
  std::uint64_t id = 0;
};

// Cumulative particle-boundary counters (simulation.h cpmlEntryCount_,
// directOuterCount_, carrierOuterCount_, MovingWindowState exits and
// croppedCarriers), with their charges.
struct BoundaryCounters {
  std::uint64_t cpml[6] = {0, 0, 0, 0, 0, 0}, direct[6] = {0, 0, 0, 0, 0, 0}, carrier[6] = {0, 0, 0, 0, 0, 0};
  std::uint64_t exits = 0, cropped = 0;
  double cpmlQ[6] = {0, 0, 0, 0, 0, 0}, directQ[6] = {0, 0, 0, 0, 0, 0}, carrierQ[6] = {0, 0, 0, 0, 0, 0};
  double exitQ = 0, croppedQ = 0;
};

struct StepResult { unsigned iterations = 0; double force = 0, work = 0; };

// Segment of one CPML carrier, processCarrierSegment (simulation.cpp:4310-
// 4407) for one rank: the carrier moves from `start` to its end position,
// is cut at the first outer-face exit, deposits its current with the
// damping at the segment midpoint and keeps the end damping.
inline void carrierSegment(Coupling& c, const ParticleBoundary& pb, Carrier carrier, const V3& start,
                           double startFraction, std::vector<Carrier>& retained, BoundaryCounters& k) {
  const Grid& g = c.fields->g;
  if (!(startFraction >= 0.0 && startFraction <= 1.0) || !std::isfinite(startFraction))
    throw std::runtime_error("Invalid CPML-carrier segment time fraction");
  if (!(carrier.weight >= 0.0) || !std::isfinite(carrier.weight))
    throw std::runtime_error("Invalid CPML carrier current weight");
  const double cutoff = 64.0 * kEpsilon;
  if (carrier.weight <= cutoff) return;
  Hit outer;
  const bool exits = pb.firstExit(start, carrier.position, outer);
  double localEnd = 1.0;
  V3 end = carrier.position;
  if (exits) {
    localEnd = outer.fraction;
    end = outer.position;
  }
  const double globalEnd = startFraction + localEnd * (1.0 - startFraction);
  const double duration = (globalEnd - startFraction) * g.dt;
  V3 mid = v3();
  for (unsigned a = 0; a < 3; ++a) mid[a] = 0.5 * (start[a] + end[a]);
  const double midDamping = pb.damping(start, mid, 0.5 * duration, g);
  const double endDamping = pb.damping(start, end, duration, g);
  const double deposited = carrier.weight * midDamping;
  const double endWeight = carrier.weight * endDamping;
  c.validatePosition(start);
  c.validatePosition(end);
  c.deposit(start, end, carrier.charge * deposited);
  if (exits) {
    ++k.carrier[outer.face];
    k.carrierQ[outer.face] += carrier.charge * endWeight;
    return;
  }
  if (endWeight <= cutoff) return;
  carrier.weight = endWeight;
  carrier.position = end;
  retained.push_back(carrier);
}

// pushDepositAndMigrate (simulation.cpp:4167-5605) for one rank: the carriers
// first (4407-4425), then every particle prepared (prepareParticle,
// 4919-5035: held particles move z += c dt beta_held with their transverse
// motion held and are released where their lab z reaches the activation
// plane, the rest of the step pushed from the release point) and its
// segment ended at the first boundary event (processActiveSegment,
// 4555-4917): CPML entry or the moving-window exit plane (the particle
// becomes a carrier from the hit), or an outer-face exit (the particle is
// removed). Held particles upstream of the window deposit only their shaped
// support (depositFrozenSourceSupport) until they cross the inlet
// (5121-5155). Retirement planes, ballistic transverse motion and the
// multi-rank transfers are not part of this path.
struct HeldRelease {
  double activationLabZ = std::numeric_limits<double>::infinity();
  // reference.hold_transverse_until_activation: z += c dt beta (true) or the
  // ballistic RelativisticBorisPusher::pushPosition (eb_particles.cpp:131-139)
  bool holdTransverse = true;
};

// pushPosition: x += (c dt / gamma) u
inline void heldBallistic(V3& x, const V3& u, double timeStep) {
  const double s = kC * timeStep / gammaFromU(u);
  for (int a = 0; a < 3; ++a) x[a] += s * u[a];
}

inline void pushDepositAll(Coupling& c, std::vector<Particle>& particles, std::vector<Carrier>& carriers,
                           BoundaryCounters& k, double time, double dt, unsigned substeps,
                           const ParticleBoundary& pb, const V3& windowOrigin, const HeldRelease& held = HeldRelease(),
                           const std::vector<Particle>* waiting = nullptr) {
  // the pending queue's particles near the inlet deposit their held
  // support first (visitNear(step, 8), simulation.cpp:4170-4172)
  if (waiting)
    for (const Particle& p : *waiting) {
      V3 end = p.position;
      end[2] += kC * dt * p.betaHeld;
      c.deposit(p.position, end, p.charge);
    }
  std::vector<Carrier> retainedCarriers;
  retainedCarriers.reserve(carriers.size());
  for (const Carrier& original : carriers) {
    Carrier carrier = original;
    const V3 start = carrier.position;
    pmv(carrier.position, dt, carrier.velocity);
    carrierSegment(c, pb, carrier, start, 0.0, retainedCarriers, k);
  }
  struct Prepared { Particle particle; bool hasActive = false; double activeStartTime = 0; Particle activeStart; };
  std::vector<Prepared> prepared(particles.size());
  for (std::size_t n = 0; n < particles.size(); ++n) {
    Prepared& r = prepared[n];
    r.particle = particles[n];
    Particle& p = r.particle;
    const V3 start = p.position;
    if (pb.window && start[2] < windowOrigin[2] && !p.frozen)
      throw std::runtime_error("An active particle is upstream of the moving window; the inlet guard did not contain its release event");
    r.activeStart = p;
    r.activeStartTime = time;
    r.hasActive = !p.frozen;
    if (p.frozen) {
      const double startLabZ = c.frame.labZFromBoxZT(start[2], time);
      if (held.holdTransverse) p.position[2] += kC * dt * p.betaHeld;
      else heldBallistic(p.position, p.u, dt);
      const double endLabZ = c.frame.labZFromBoxZT(p.position[2], time + dt);
      if (endLabZ >= held.activationLabZ) {
        if (pb.window && start[2] < windowOrigin[2])
          throw std::runtime_error("A pending upstream held particle would release before a complete in-window step; check the moving inlet guard");
        const double f = startLabZ >= held.activationLabZ
                             ? 0.0
                             : std::max(0.0, std::min(1.0, (held.activationLabZ - startLabZ) / (endLabZ - startLabZ)));
        p.position = start;
        if (held.holdTransverse) p.position[2] += kC * f * dt * p.betaHeld;
        else heldBallistic(p.position, p.u, f * dt);
        p.frozen = false;
        r.hasActive = true;
        r.activeStart = p;
        r.activeStartTime = time + f * dt;
        const double remaining = (1.0 - f) * dt;
        if (remaining > 0.0) c.push(p, r.activeStartTime, remaining, substeps);
      }
    } else {
      c.push(p, time, dt, substeps);
    }
  }
  const double tolerance = 64.0 * kEpsilon;
  std::vector<Particle> retained;
  retained.reserve(prepared.size());
  for (std::size_t n = 0; n < particles.size(); ++n) {
    V3 start = particles[n].position;
    double startFraction = 0.0;
    const Particle& p = prepared[n].particle;
    if (pb.window && start[2] < windowOrigin[2]) {
      if (!particles[n].frozen || !p.frozen || prepared[n].hasActive)
        throw std::runtime_error("Moving-window upstream admission requires an unreleased held particle on rank zero");
      if (p.position[2] < windowOrigin[2]) {
        c.deposit(start, p.position, p.charge);  // depositFrozenSourceSupport
        retained.push_back(p);
        continue;
      }
      const double fraction = (windowOrigin[2] - start[2]) / (p.position[2] - start[2]);
      if (!(fraction >= 0.0 && fraction <= 1.0) || !std::isfinite(fraction))
        throw std::runtime_error("Invalid moving-window admission fraction");
      for (unsigned a = 0; a < 3; ++a) start[a] += fraction * (p.position[a] - start[a]);
      start[2] = windowOrigin[2];
      startFraction = fraction;
      c.deposit(particles[n].position, start, p.charge);  // the support's prefix
    }
    const bool kink = particles[n].frozen && prepared[n].hasActive && prepared[n].activeStartTime > time &&
                      prepared[n].activeStartTime < time + dt;
    if (kink) {
      // processActiveSegment's frozen leg (4676-4746)
      const V3& knot = prepared[n].activeStart.position;
      const double kinkFraction = (prepared[n].activeStartTime - time) / dt;
      Hit h1, h2;
      if (pb.firstEntry(start, knot, h1) || pb.firstExit(start, knot, h2))
        throw std::runtime_error("Held-transverse activation crossed a particle boundary during the frozen part of a field step; enable mesh.ballistic_outside_transverse for boundary injection");
      c.validatePosition(start);
      c.validatePosition(knot);
      c.deposit(start, knot, p.charge);
      start = knot;
      startFraction = kinkFraction;
    }
    Hit cpmlHit, outerHit;
    const bool entersCpml = pb.window ? pb.windowEntry(start, p.position, time + startFraction * dt, time + dt, c.frame, cpmlHit)
                                      : pb.firstEntry(start, p.position, cpmlHit);
    const bool exitsOuter = pb.firstExit(start, p.position, outerHit);
    enum { Retain, EnterCpml, ExitOuter } action = Retain;
    double localEnd = 1.0;
    V3 end = p.position;
    Hit selected;
    if (entersCpml && (!exitsOuter || cpmlHit.fraction <= outerHit.fraction + tolerance)) {
      action = EnterCpml;
      localEnd = cpmlHit.fraction;
      end = cpmlHit.position;
      selected = cpmlHit;
    } else if (exitsOuter) {
      action = ExitOuter;
      localEnd = outerHit.fraction;
      end = outerHit.position;
      selected = outerHit;
    }
    c.validatePosition(start);
    c.validatePosition(end);
    if (!std::isfinite(p.charge)) throw std::invalid_argument("Deposited particle charge must be finite");
    c.deposit(start, end, p.charge);
    const double globalEnd = startFraction + localEnd * (1.0 - startFraction);
    if (action == EnterCpml) {
      if (pb.window && selected.face == 5) {
        ++k.exits;
        k.exitQ += p.charge;
      } else {
        ++k.cpml[selected.face];
        k.cpmlQ[selected.face] += p.charge;
      }
      Carrier carrier;
      carrier.position = p.position;
      const double gamma = gammaFromU(p.u);
      carrier.velocity = p.u;
      scaleBy(carrier.velocity, kC / gamma);
      carrier.u = p.u;
      carrier.charge = p.charge;
      carrier.weight = 1.0;
      carrier.id = p.id;
      carrierSegment(c, pb, carrier, selected.position, globalEnd, retainedCarriers, k);
    } else if (action == ExitOuter) {
      ++k.direct[selected.face];
      k.directQ[selected.face] += p.charge;
    } else {
      retained.push_back(p);
    }
  }
  particles.swap(retained);
  carriers.swap(retainedCarriers);
}

// The carrier part of Simulation::finishMovingWindowStep after the shift
// (simulation_moving_window.cpp:353-367): active particles must not lie
// upstream of the new origin; carriers at or beyond the new upper face are
// cropped and counted.
inline void finishWindowStep(const Grid& g, const V3& newOrigin, std::vector<Particle>& particles,
                             std::vector<Carrier>& carriers, BoundaryCounters& k) {
  for (const auto& p : particles) {
    if (p.position[2] < newOrigin[2] && !p.frozen) throw std::runtime_error("Active particle fell behind moving inlet");
    // migrateWindowObjects with ownerRank (simulation.cpp:4068-4083) on one
    // rank: a particle beyond the upper face has no owner.
    const double coordinate = (p.position[2] - newOrigin[2]) / g.dz;
    const double tolerance = 64.0 * kEpsilon * std::max(1.0, std::abs(coordinate));
    if (coordinate > static_cast<double>(g.nz) + tolerance)
      throw std::runtime_error("Moving-window remap particle crossed a non-adjacent slab or outer boundary");
  }
  const double upper = std::fma(static_cast<double>(g.nz), g.dz, newOrigin[2]);
  std::size_t retained = 0;
  for (const auto& q : carriers) {
    if (q.position[2] >= upper) {
      ++k.cropped;
      k.croppedQ += q.charge * q.weight;
    } else if (q.position[2] < newOrigin[2]) {
      throw std::runtime_error("Current carrier fell behind moving inlet");
    } else {
      carriers[retained++] = q;
    }
  }
  carriers.resize(retained);
}

// TransverseReferenceCurrent (transverse_reference_current.cpp:18-26,
// 104-120) without its file cache: the prepared surface current of the
// reference field at the four transverse interfaces of the aperture, one
// table row per reference z cell, added to Jz every trial and at the
// commit, interpolated linearly at z = offset + k - beta (step + 1).
struct RefCurrent {
  struct Point { std::size_t i, j; };
  std::vector<Point> points;
  std::vector<double> table;  // rows x points
  std::size_t rows = 0;
  double beta = 0.0;
  bool enabled = false;

  // points(), transverse_reference_current.cpp:18-26
  void setPoints(const Grid& g, std::size_t gx, std::size_t gy) {
    if (!gx || !gy) throw std::invalid_argument("Reference surface current requires transverse-only CPML");
    points.clear();
    const std::size_t x0 = gx, x1 = g.nx - gx, y0 = gy, y1 = g.ny - gy;
    for (std::size_t i = x0; i <= x1; ++i)
      for (std::size_t j : {y0, y1}) points.push_back({i, j});
    for (std::size_t j = y0 + 1; j < y1; ++j)
      for (std::size_t i : {x0, x1}) points.push_back({i, j});
  }
  // apply, transverse_reference_current.cpp:104-120 (no work audit)
  void apply(Fields& f, std::ptrdiff_t offset, std::size_t step) const {
    const std::size_t nz = f.g.nz, count = points.size();
    Comp& jz = f.f[8];
    for (std::size_t k = 0; k < nz; ++k) {
      const double z = static_cast<double>(offset) + k - beta * static_cast<double>(step + 1);
      if (!std::isfinite(z) || z < 0. || z > rows - 1) throw std::out_of_range("Reference surface-current query outside domain");
      const auto lo = static_cast<std::size_t>(std::floor(z)), hi = std::min<std::size_t>(lo + 1, rows - 1);
      const double fraction = z - lo;
      const double* a = table.data() + lo * count;
      const double* b = table.data() + hi * count;
      for (std::size_t p = 0; p < count; ++p) {
        const double value = fraction == 0. ? a[p] : (1. - fraction) * a[p] + fraction * b[p];
        jz(points[p].i, points[p].j, k) += value;
      }
    }
  }
};

// ChargeMatchedInlet (charge_matched_inlet.h:16-72) for one rank: after a
// shift, the new inlet layer's Ez on the interior aperture is set so that
// div E at the inlet vertices matches the shaped charge of the particles
// near the inlet: Ez(k=0) = Ez(k=1) + dz (div_xy E(k=1) - rho / eps0), in
// long double like the reference. origin = the shifted window's origin.
inline void chargeMatchedInlet(Fields& f, const std::vector<Particle>& particles, const V3& origin, int shape,
                               bool xBinomial, std::size_t gx, std::size_t gy) {
  const Grid& g = f.g;
  if (g.nz < 2 || gx + 1 >= g.nx - gx || gy + 1 >= g.ny - gy)
    throw std::invalid_argument("Charge-matched inlet needs an interior physical aperture");
  std::vector<double> rho((g.nx + 1) * (g.ny + 1), 0.0);
  for (const Particle& p : particles) {
    const double z = (p.position[2] - origin[2]) / g.dz;
    const Basis b = basis(z, false, shape);
    const double wz = 1 < b.first || 1 >= b.first + static_cast<long long>(b.count) ? 0. : b.v[1 - b.first];
    if (wz == 0.) continue;
    const double x = (p.position[0] - origin[0]) / g.dx, y = (p.position[1] - origin[1]) / g.dy;
    if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z) || !std::isfinite(p.charge))
      throw std::runtime_error("Nonfinite charge-matched inlet source");
    if (y < 0. || y > static_cast<double>(g.ny)) continue;
    if (xBinomial ? (x <= -2. || x >= static_cast<double>(g.nx) + 2.) : (x < 0. || x > static_cast<double>(g.nx)))
      continue;
    const long long ix = static_cast<long long>(std::floor(x));
    const std::size_t j = static_cast<std::size_t>(std::floor(y));
    const double fx = x - ix, fy = y - j, scale = p.charge * (1. / (g.dx * g.dy * g.dz));
    const std::array<double, 4> wx = xBinomial ? nodeWeights(fx) : std::array<double, 4>{{1. - fx, fx, 0., 0.}};
    const unsigned nX = xBinomial ? 4 : 2;
    const long long first = ix - (xBinomial ? 1 : 0);
    for (unsigned bb = 0; bb < 2; ++bb)
      for (unsigned a = 0; a < nX; ++a) {
        const long long i = first + a;
        if (i >= 0 && i <= static_cast<long long>(g.nx) && j + bb <= g.ny)
          rho[(j + bb) * (g.nx + 1) + static_cast<std::size_t>(i)] += scale * wx[a] * (bb ? fy : 1. - fy) * wz;
      }
  }
  Comp& ex = f.f[0];
  Comp& ey = f.f[1];
  Comp& ez = f.f[2];
  for (std::size_t j = gy + 1; j < g.ny - gy; ++j)
    for (std::size_t i = gx + 1; i < g.nx - gx; ++i) {
      const long double divxy = (static_cast<long double>(ex(i, j, 1)) - ex(i - 1, j, 1)) / g.dx +
                                (static_cast<long double>(ey(i, j, 1)) - ey(i, j - 1, 1)) / g.dy;
      const long double r = rho[j * (g.nx + 1) + i], retained = ez(i, j, 1);
      const double value = static_cast<double>(retained + g.dz * (divxy - r / kEps0));
      if (!std::isfinite(value)) throw std::runtime_error("Nonfinite charge-matched inlet field");
      ez(i, j, 0) = value;
    }
}

// Simulation::solveConservativeSourceStep + commitConservativeElectric
// (simulation_source_coupling.cpp:16-65); the particles, the carriers and the
// boundary counters are restored before every trial (restore(), 24-31).
inline StepResult sourceStep(Fields& f, Coupling& c, std::vector<Particle>& particles, std::vector<Carrier>& carriers,
                             BoundaryCounters& counters, const V3& origin, double time, unsigned substeps,
                             const ParticleBoundary& pb, const RefCurrent* ref = nullptr,
                             const HeldRelease& held = HeldRelease(), const std::vector<Particle>* waiting = nullptr) {
  c.begin(f, origin, particles);
  f.window->electricPreview = true;
  f.advanceElectric();
  f.window->electricPreview = false;
  const std::vector<Particle> original = particles;
  const std::vector<Carrier> originalCarriers = carriers;
  const BoundaryCounters originalCounters = counters;
  bool converged = false;
  StepResult r;
  const unsigned maximum = 64;
  for (unsigned n = 0; n < maximum; ++n) {
    if (n) c.applyCurrentSource(1.);
    f.clearCurrent();
    particles = original;
    carriers = originalCarriers;
    counters = originalCounters;
    c.beginTrial();
    pushDepositAll(c, particles, carriers, counters, time, f.g.dt, substeps, pb, origin, held, waiting);
    // applyMovingTransverseReferenceCurrent (simulation_source_coupling.cpp:44)
    if (ref && ref->enabled) ref->apply(f, f.window->offset, f.window->step);
    c.applyCurrentSource(-1.);
    const auto res = c.localResidual();
    r.force = res[0];
    const double total[3] = {res[1], res[2], res[3]};
    r.work = total[2] == 0. ? 0. : (total[1] > 0. ? total[0] / total[1] : total[0]);
    r.iterations = n + 1;
    if (r.iterations >= (total[2] > 0. ? 2u : 1u) && r.force < 1e-13 && r.work < 1e-8) {
      converged = true;
      break;
    }
  }
  if (!converged) throw std::runtime_error("Conservative source coupling failed to converge");
  c.applyCurrentSource(1.);
  f.clearCurrent();
  particles = original;
  carriers = originalCarriers;
  counters = originalCounters;
  c.swapElectric();
  c.replay = true;
  pushDepositAll(c, particles, carriers, counters, time, f.g.dt, substeps, pb, origin, held, waiting);
  c.replay = false;
  // the solve loop's reference current before the commit (simulation.cpp:548-549)
  if (ref && ref->enabled) ref->apply(f, f.window->offset, f.window->step);
  // commitConservativeElectric
  c.swapElectric();
  f.undoPreview();
  c.applyCurrentSource(-1.);
  if (f.window->electricPreview || f.window->previewPresent) throw std::logic_error("Electric preview not restored before commit");
  f.afterElectricBoundary();
  return r;
}

}  // namespace orc
