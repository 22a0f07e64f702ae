// Templates of the field kernels and their host launchers for a storage
// type R (double: the plain form; float: the FP32 form), included by the
// translation units that instantiate them (fields.hip for R = double,
// fields32.hip for R = float). Updates are evaluated in FP64 for both and
// rounded once to R on store.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "fields.hpp"
#include "tilecounts.hpp"
#include "tuner.hpp"

namespace aprl_gpu {

template <class R>
void allocateFields(FieldStateT<R>& s, std::size_t slack) {
  const std::size_t cells[3] = {s.g.nx, s.g.ny, s.g.nz};
  for (int a = 0; a < 9; ++a) {
    const bool electric = a < 3 || a >= 6;
    const int axis = a % 3;
    for (int d = 0; d < 3; ++d) s.dims[a][d] = cells[d] + (electric ? (d != axis) : (d == axis));
  }
  s.slack = slack;
  s.zb = slack;
  std::size_t largest = 0;
  for (int a = 0; a < 9; ++a) {
    const std::size_t planes = s.dims[a][2] + (s.slides(a) ? slack : 0);
    s.comp[a].resize(planes * s.plane(a));
    hipCheck(hipMemsetAsync(s.comp[a].get(), 0, s.comp[a].size() * sizeof(R), s.stream), "hipMemsetAsync");
    largest = std::max(largest, s.compSize(a));
  }
  if (slack) s.scratch.resize(largest);
}

template <class R>
FieldPtrsT<R> FieldStateT<R>::ptrs() const {
  FieldPtrsT<R> f{};
  for (int a = 0; a < 3; ++a) {
    f.e[a] = base(a);
    f.b[a] = base(3 + a);
    f.j[a] = base(6 + a);
    f.lo[a] = nullptr;
    f.up[a] = nullptr;
  }
  for (int a = 0; a < 9; ++a)
    for (int d = 0; d < 3; ++d) f.n[a][d] = static_cast<int>(dims[a][d]);
  return f;
}

// Coefficients stay in double for both storage types: the FP32 form stores
// its fields in FP32 but evaluates every update in FP64 and rounds the
// result once, so no coefficient rounding biases its dispersion.
template <class R>
StepCoefsT<R> FieldStateT<R>::coefs() const {
  StepCoefsT<R> c{};
  for (int a = 0; a < 3; ++a) { c.alpha[a] = cw.alpha[a]; c.beta[a] = cw.beta[a]; }
  c.dYZ = cw.deltaYZ; c.dZX = cw.deltaZX; c.dXY = cw.deltaXY;
  c.tx = g.dt / g.dx; c.ty = g.dt / g.dy; c.tz = g.dt / g.dz;
  const double curlScale = g.dt / (kMu0 * kEps0);
  c.cX = curlScale / g.dx; c.cY = curlScale / g.dy;
  c.cZ = curlScale / g.dz;
  c.cs = g.dt / kEps0;
  return c;
}

// Pair order of the checkpoint histories, eb_cpml.cpp:355-357 / 556-558.
static const int kPairs[6][2] = {{0, 1}, {0, 2}, {1, 2}, {1, 0}, {2, 0}, {2, 1}};

template <class R>
CpmlArgsT<R> FieldStateT<R>::cpmlArgs() const {
  CpmlArgsT<R> P{};
  P.enabled = hasCpml ? 1 : 0;
  P.lowerZ = lowerZ ? 1 : 0;
  P.upperZ = upperZ ? 1 : 0;
  P.zb = static_cast<int>(zb);
  if (!hasCpml) return P;
  for (int m = 0; m < 2; ++m)
    for (int a = 0; a < 3; ++a) { P.prof[m][a] = prof[m][a].get(); P.ord[m][a] = ord[m][a].get(); }
  for (int h = 0; h < 12; ++h) {
    const int m = h < 6 ? 0 : 1, t = kPairs[h % 6][0], d = kPairs[h % 6][1];
    P.hist[m][t][d] = hist[h].get();
    P.row[m][t][d] = static_cast<int>(histRow[h]);
    P.ostride[m][t][d] = static_cast<long long>(histSlides[h] ? (histLayers[h] + slack) * histRow[h]
                                                               : histLayers[h] * histRow[h]);
  }
  const double curlScale = g.dt / (kMu0 * kEps0);
  P.magScale[0] = g.dt / g.dx; P.magScale[1] = g.dt / g.dy;
  P.magScale[2] = g.dt / g.dz;
  P.eleScale[0] = curlScale / g.dx; P.eleScale[1] = curlScale / g.dy;
  P.eleScale[2] = curlScale / g.dz;
  return P;
}

// Offset of a history slot's first stored layer (checkpoint layer 0).
template <class R>
std::size_t FieldStateT<R>::historyDeviceOffset(std::size_t h, std::size_t slot) const {
  if (!histSlides[h]) return slot * histLayers[h] * histRow[h];
  return slot * (histLayers[h] + slack) * histRow[h] + zb * histRow[h];
}

// EBCPMLParameters::maximumConductivityRate, eb_cpml.cpp:263-281.
static double maxRate(const aprl_cpml_parameters& p, int axis, const aprl_grid_geometry& g) {
  if (p.cells[axis] == 0) return 0.0;
  const double spacing = axis == 0 ? g.dx : (axis == 1 ? g.dy : g.dz);
  if (!(spacing > 0.0) || !std::isfinite(spacing) || !(p.polynomial_order > 0.0) ||
      !(p.target_reflection > 0.0 && p.target_reflection < 1.0))
    throw std::invalid_argument("Cannot evaluate an invalid CPML conductivity profile");
  const double thickness = static_cast<double>(p.cells[axis]) * spacing;
  return -(p.polynomial_order + 1.0) * kC * std::log(p.target_reflection) / (2.0 * thickness);
}

// EBConvolutionalPML construction, eb_cpml.cpp:195-248 and 297-372; the
// profiles are evaluated on the host with the same libm calls.
template <class R>
void installCpml(FieldStateT<R>& s, const aprl_cpml_parameters& p, std::size_t globalNz, std::size_t zOffset,
                 bool lowerZ, bool upperZ) {
  if (!(p.cells[0] > 0 || p.cells[1] > 0 || p.cells[2] > 0))
    throw std::invalid_argument("CPML requires at least one nonzero boundary thickness");
  if (!(p.polynomial_order > 0.0) || !(p.target_reflection > 0.0 && p.target_reflection < 1.0) ||
      !(p.kappa_max >= 1.0) || !(p.alpha_fraction >= 0.0) || !std::isfinite(p.polynomial_order) ||
      !std::isfinite(p.target_reflection) || !std::isfinite(p.kappa_max) || !std::isfinite(p.alpha_fraction))
    throw std::invalid_argument("Invalid CPML grading parameters");
  const aprl_grid_geometry& g = s.g;
  const std::size_t gc[3] = {g.nx, g.ny, globalNz};
  for (int a = 0; a < 3; ++a)
    if (p.cells[a] > 0) {
      if (p.cells[a] < 2) throw std::invalid_argument("Each enabled CPML direction needs at least two cells");
      if (2 * p.cells[a] >= gc[a]) throw std::invalid_argument("Opposite CPML slabs must leave a non-PML interior");
    }
  if (p.cells[2] > 0) {
    if (!lowerZ && zOffset < p.cells[2])
      throw std::invalid_argument("The lower z CPML slab cannot span MPI ranks yet; reduce MPI ranks or CPML cells");
    if (!upperZ && zOffset + g.nz > globalNz - p.cells[2])
      throw std::invalid_argument("The upper z CPML slab cannot span MPI ranks yet; reduce MPI ranks or CPML cells");
  }
  if (s.slack && p.cells[2] > 0) throw std::invalid_argument("Native-z history transport requires transverse-only CPML");
  std::size_t activeCount[2][3] = {};
  for (int m = 0; m < 2; ++m) {
    const bool electric = m == 1;
    for (int axis = 0; axis < 3; ++axis) {
      const std::size_t localCells = axis == 0 ? g.nx : (axis == 1 ? g.ny : g.nz);
      const std::size_t globalCells = axis == 2 ? globalNz : localCells;
      const std::size_t globalOffset = axis == 2 ? zOffset : 0;
      const std::size_t count = localCells + (electric ? 1 : 0);
      std::vector<CpmlPointDev> pts(count, CpmlPointDev{0.0, 0.0, 0.0});
      std::vector<int> ordv(count, -1);
      const std::size_t thicknessCells = p.cells[axis];
      if (thicknessCells > 0) {
        const double thickness = static_cast<double>(thicknessCells);
        const double smax = maxRate(p, axis, g);
        int ordinal = 0;
        for (std::size_t li = 0; li < count; ++li) {
          const double position = static_cast<double>(globalOffset + li) + (electric ? 0.0 : 0.5);
          double depth = 0.0;
          if (position < thickness) depth = (thickness - position) / thickness;
          const double upperStart = static_cast<double>(globalCells - thicknessCells);
          if (position > upperStart) depth = std::max(depth, (position - upperStart) / thickness);
          if (!(depth > 0.0)) continue;
          depth = std::min(1.0, depth);
          const double graded = std::pow(depth, p.polynomial_order);
          const double sigma = smax * graded;
          const double kappa = 1.0 + (p.kappa_max - 1.0) * graded;
          const double alpha = p.alpha_fraction * smax * (1.0 - depth);
          CpmlPointDev& c = pts[li];
          c.ikm1 = 1.0 / kappa - 1.0;
          c.decay = std::exp(-(sigma / kappa + alpha) * g.dt);
          const double den = sigma * kappa + alpha * kappa * kappa;
          c.drive = den > 0.0 ? sigma * (c.decay - 1.0) / den : 0.0;
          ordv[li] = ordinal++;
        }
      }
      activeCount[m][axis] = static_cast<std::size_t>(std::count_if(ordv.begin(), ordv.end(), [](int o) { return o >= 0; }));
      std::vector<CpmlPointT<R>> ptsR(count);
      for (std::size_t li = 0; li < count; ++li)
        ptsR[li] = CpmlPointT<R>{pts[li].ikm1, pts[li].decay, pts[li].drive};
      s.prof[m][axis].resize(count);
      s.ord[m][axis].resize(count);
      hipCheck(hipMemcpy(s.prof[m][axis].get(), ptsR.data(), count * sizeof(CpmlPointT<R>), hipMemcpyHostToDevice),
               "hipMemcpy");
      hipCheck(hipMemcpy(s.ord[m][axis].get(), ordv.data(), count * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
    }
  }
  for (int h = 0; h < 12; ++h) {
    const int m = h < 6 ? 0 : 1, t = kPairs[h % 6][0], d = kPairs[h % 6][1];
    const int a = m == 0 ? 3 + t : t;
    const std::size_t* dim = s.dims[a];
    s.histSlots[h] = activeCount[m][d];
    if (d == 2) {
      s.histSlides[h] = false;
      s.histLayers[h] = dim[1];
      s.histRow[h] = dim[0];
    } else {
      s.histSlides[h] = s.slack > 0;
      s.histLayers[h] = dim[2];
      s.histRow[h] = d == 0 ? dim[1] : dim[0];
    }
    const std::size_t layers = s.histLayers[h] + (s.histSlides[h] ? s.slack : 0);
    s.hist[h].resize(s.histSlots[h] * layers * s.histRow[h]);
    if (s.hist[h].size())
      hipCheck(hipMemsetAsync(s.hist[h].get(), 0, s.hist[h].size() * sizeof(R), s.stream), "hipMemsetAsync");
  }
  s.cpml = p;
  s.hasCpml = true;
  s.lowerZ = lowerZ;
  s.upperZ = upperZ;
}

// ---------------------------------------------------------------------------
// Device code.

// EBFieldGrid::electricValue, eb_field.cpp:350-417.
template <int A, class R>
__device__ __forceinline__ double evalE(const FieldPtrsT<R>& F, int i, int j, int k) {
  const int n0 = F.n[A][0], n1 = F.n[A][1], n2 = F.n[A][2];
  double sign = 1.0;
  if (i < 0) {
    if (A == 0) i = 0;
    else { i = -i; sign = -sign; }
  } else if (i >= n0) {
    if (A == 0) i = n0 - 1;
    else { i = 2 * (n0 - 1) - i; sign = -sign; }
  }
  if (j < 0) {
    if (A == 1) j = 0;
    else { j = -j; sign = -sign; }
  } else if (j >= n1) {
    if (A == 1) j = n1 - 1;
    else { j = 2 * (n1 - 1) - j; sign = -sign; }
  }
  if (k < 0 || k >= n2) {
    const bool lower = k < 0;
    const R* plane = lower ? F.lo[A] : F.up[A];
    if (plane) return sign * static_cast<double>(plane[static_cast<size_t>(j) * n0 + i]);
    if (A == 2) k = lower ? 0 : n2 - 1;
    else { k = lower ? -k : 2 * (n2 - 1) - k; sign = -sign; }
  }
  return sign * static_cast<double>(F.e[A][(static_cast<size_t>(k) * n1 + j) * n0 + i]);
}

// EBFieldGrid::smoothedElectric, eb_field.cpp:419-486 (same accumulation
// order on the interior fast path and the boundary path).
template <int A, int D, class R>
__device__ __forceinline__ double smoothE(const FieldPtrsT<R>& F, const StepCoefsT<R>& c, int i, int j, int k) {
  constexpr int TA = (D + 1) % 3, TB = (D + 2) % 3;
  const double delta = D == 0 ? c.dYZ : (D == 1 ? c.dZX : c.dXY);
  const int n0 = F.n[A][0], n1 = F.n[A][1], n2 = F.n[A][2];
  if (i > 0 && i < n0 - 1 && j > 0 && j < n1 - 1 && k > 0 && k < n2 - 1) {
    const long long st[3] = {1, n0, static_cast<long long>(n0) * n1};
    const long long oA = st[TA], oB = st[TB];
    const R* p = F.e[A] + (static_cast<size_t>(k) * n1 + j) * n0 + i;
    double r = c.alpha[D] * static_cast<double>(p[0]);
    r += c.beta[TA] * static_cast<double>(p[-oA]);
    r += c.beta[TB] * static_cast<double>(p[-oB]);
    r += c.beta[TA] * static_cast<double>(p[oA]);
    r += c.beta[TB] * static_cast<double>(p[oB]);
    r += delta * static_cast<double>(p[-oA - oB]);
    r += delta * static_cast<double>(p[-oA + oB]);
    r += delta * static_cast<double>(p[oA - oB]);
    r += delta * static_cast<double>(p[oA + oB]);
    return r;
  }
  int b[3] = {i, j, k};
  double r = c.alpha[D] * evalE<A>(F, i, j, k);
  for (int dir = -1; dir <= 1; dir += 2) {
    int pA[3] = {b[0], b[1], b[2]}, pB[3] = {b[0], b[1], b[2]};
    pA[TA] += dir;
    pB[TB] += dir;
    r += c.beta[TA] * evalE<A>(F, pA[0], pA[1], pA[2]);
    r += c.beta[TB] * evalE<A>(F, pB[0], pB[1], pB[2]);
  }
  for (int dA = -1; dA <= 1; dA += 2)
    for (int dB = -1; dB <= 1; dB += 2) {
      int p[3] = {b[0], b[1], b[2]};
      p[TA] += dA;
      p[TB] += dB;
      r += delta * evalE<A>(F, p[0], p[1], p[2]);
    }
  return r;
}

template <int A, class R>
__device__ __forceinline__ double rawE(const FieldPtrsT<R>& F, int i, int j, int k) {
  return static_cast<double>(F.e[A][(static_cast<size_t>(k) * F.n[A][1] + j) * F.n[A][0] + i]);
}

// Index of a CPML history entry of target (M, T) and derivative D at node
// (i, j, k), CPMLMemoryTerm::value (eb_cpml.cpp:78-89) with slid layers.
template <int M, int T, int D, class R>
__device__ __forceinline__ R* histAt(const CpmlArgsT<R>& P, int o, int i, int j, int k, int dim0) {
  R* h = P.hist[M][T][D] + static_cast<long long>(o) * P.ostride[M][T][D];
  if (D == 0) return h + static_cast<long long>(P.zb + k) * P.row[M][T][D] + j;
  if (D == 1) return h + static_cast<long long>(P.zb + k) * P.row[M][T][D] + i;
  return h + static_cast<long long>(j) * dim0 + i;
}

// transformedCorrection, eb_cpml.cpp:131-138. WRITE selects whether the
// advanced history is stored (preview keeps the old one).
template <bool WRITE, class R>
__device__ __forceinline__ double cpmlCorrection(const CpmlPointT<R>& c, double raw, R* hist, R old) {
  const double h = c.decay * static_cast<double>(old) + c.drive * raw;
  if (WRITE) *hist = static_cast<R>(h);
  return c.ikm1 * raw + h;
}
template <bool WRITE, class R>
__device__ __forceinline__ double cpmlCorrection(const CpmlPointT<R>& c, double raw, R* hist) {
  return cpmlCorrection<WRITE>(c, raw, hist, *hist);
}

// Per-component magnetic updates at one node: advanceMagneticCowan
// (eb_field.cpp:488-558) / advanceMagneticYee (311-348) followed by
// EBConvolutionalPML::afterMagneticUpdate (406-468). dA/dB are the two curl
// differences of the component; the CPML terms use the same differences.
template <class R>
__device__ __forceinline__ size_t bIndex(const FieldPtrsT<R>& F, int a, int i, int j, int k) {
  return (static_cast<size_t>(k) * F.n[3 + a][1] + j) * F.n[3 + a][0] + i;
}
// The loads of component A's update at node (i, j, k): its old value and
// the CPML histories of its two terms (null pointer: term inactive). oI, oJ,
// oK are the magnetic CPML ordinals of i, j, k (-1 outside the layers or
// without CPML). Term order per component (afterMagneticUpdate): Bx (y, z),
// By (z, x), Bz (x, y).
template <class R>
struct BLoads {
  R old;
  R h[2];
  R* hp[2];
};
template <int A, class R>
__device__ __forceinline__ BLoads<R> loadB(const FieldPtrsT<R>& F, const CpmlArgsT<R>& P, int i, int j, int k, int oI,
                                           int oJ, int oK) {
  const int nx = F.n[0][0];
  BLoads<R> L;
  L.old = F.b[A][bIndex(F, A, i, j, k)];
  const int o1 = A == 0 ? oJ : (A == 1 ? oK : oI), o2 = A == 0 ? oK : (A == 1 ? oI : oJ);
  L.hp[0] = nullptr;
  L.hp[1] = nullptr;
  L.h[0] = 0;
  L.h[1] = 0;
  if (o1 >= 0) {
    L.hp[0] = A == 0 ? histAt<0, 0, 1>(P, o1, i, j, k, nx + 1)
                     : (A == 1 ? histAt<0, 1, 2>(P, o1, i, j, k, nx) : histAt<0, 2, 0>(P, o1, i, j, k, nx));
    L.h[0] = *L.hp[0];
  }
  if (o2 >= 0) {
    L.hp[1] = A == 0 ? histAt<0, 0, 2>(P, o2, i, j, k, nx + 1)
                     : (A == 1 ? histAt<0, 1, 0>(P, o2, i, j, k, nx) : histAt<0, 2, 1>(P, o2, i, j, k, nx));
    L.h[1] = *L.hp[1];
  }
  return L;
}
// The update itself: advanceMagneticCowan / advanceMagneticYee's curl step
// with the two curl differences d1, d2 (Bx: dEz/dy, dEy/dz; By: dEx/dz,
// dEz/dx; Bz: dEy/dx, dEx/dy), then the CPML terms, then the stores.
template <int A, class R>
__device__ __forceinline__ void finishB(FieldPtrsT<R>& F, const StepCoefsT<R>& c, const CpmlArgsT<R>& P, int i, int j,
                                        int k, double d1, double d2, const BLoads<R>& L) {
  constexpr int S1 = A == 0 ? 1 : (A == 1 ? 2 : 0), S2 = A == 0 ? 2 : (A == 1 ? 0 : 1);
  const double t1 = A == 0 ? c.ty : (A == 1 ? c.tz : c.tx), t2 = A == 0 ? c.tz : (A == 1 ? c.tx : c.ty);
  const int ax1 = S1, ax2 = S2;
  double b = static_cast<double>(L.old);
  b -= t1 * d1 - t2 * d2;
  if (L.hp[0]) {
    const int at1 = ax1 == 0 ? i : (ax1 == 1 ? j : k);
    b += (-1.0 * P.magScale[S1]) * cpmlCorrection<true>(P.prof[0][ax1][at1], d1, L.hp[0], L.h[0]);
  }
  if (L.hp[1]) {
    const int at2 = ax2 == 0 ? i : (ax2 == 1 ? j : k);
    b += (1.0 * P.magScale[S2]) * cpmlCorrection<true>(P.prof[0][ax2][at2], d2, L.hp[1], L.h[1]);
  }
  F.b[A][bIndex(F, A, i, j, k)] = static_cast<R>(b);
}
template <class R>
__device__ __forceinline__ void finishBx(FieldPtrsT<R>& F, const StepCoefsT<R>& c, const CpmlArgsT<R>& P, int i, int j, int k,
                                         double dEzDy, double dEyDz) {
  const int oJ = P.enabled ? P.ord[0][1][j] : -1, oK = P.enabled ? P.ord[0][2][k] : -1;
  finishB<0>(F, c, P, i, j, k, dEzDy, dEyDz, loadB<0>(F, P, i, j, k, -1, oJ, oK));
}
template <class R>
__device__ __forceinline__ void finishBy(FieldPtrsT<R>& F, const StepCoefsT<R>& c, const CpmlArgsT<R>& P, int i, int j, int k,
                                         double dExDz, double dEzDx) {
  const int oK = P.enabled ? P.ord[0][2][k] : -1, oI = P.enabled ? P.ord[0][0][i] : -1;
  finishB<1>(F, c, P, i, j, k, dExDz, dEzDx, loadB<1>(F, P, i, j, k, oI, -1, oK));
}
template <class R>
__device__ __forceinline__ void finishBz(FieldPtrsT<R>& F, const StepCoefsT<R>& c, const CpmlArgsT<R>& P, int i, int j, int k,
                                         double dEyDx, double dExDy) {
  const int oI = P.enabled ? P.ord[0][0][i] : -1, oJ = P.enabled ? P.ord[0][1][j] : -1;
  finishB<2>(F, c, P, i, j, k, dEyDx, dExDy, loadB<2>(F, P, i, j, k, oI, oJ, -1));
}

// One thread per node over the union of the three B extents (Yee; and the
// Cowan edge nodes Bx(i = nx), By(j = ny) when EDGES).
template <bool COWAN, class R>
__device__ __forceinline__ void magneticBx(FieldPtrsT<R>& F, const StepCoefsT<R>& c, const CpmlArgsT<R>& P, int i, int j, int k) {
  double dEzDy, dEyDz;
  if (COWAN) {
    dEzDy = smoothE<2, 1>(F, c, i, j + 1, k) - smoothE<2, 1>(F, c, i, j, k);
    dEyDz = smoothE<1, 2>(F, c, i, j, k + 1) - smoothE<1, 2>(F, c, i, j, k);
  } else {
    dEzDy = rawE<2>(F, i, j + 1, k) - rawE<2>(F, i, j, k);
    dEyDz = rawE<1>(F, i, j, k + 1) - rawE<1>(F, i, j, k);
  }
  finishBx(F, c, P, i, j, k, dEzDy, dEyDz);
}
template <bool COWAN, class R>
__device__ __forceinline__ void magneticBy(FieldPtrsT<R>& F, const StepCoefsT<R>& c, const CpmlArgsT<R>& P, int i, int j, int k) {
  double dExDz, dEzDx;
  if (COWAN) {
    dExDz = smoothE<0, 2>(F, c, i, j, k + 1) - smoothE<0, 2>(F, c, i, j, k);
    dEzDx = smoothE<2, 0>(F, c, i + 1, j, k) - smoothE<2, 0>(F, c, i, j, k);
  } else {
    dExDz = rawE<0>(F, i, j, k + 1) - rawE<0>(F, i, j, k);
    dEzDx = rawE<2>(F, i + 1, j, k) - rawE<2>(F, i, j, k);
  }
  finishBy(F, c, P, i, j, k, dExDz, dEzDx);
}

template <class R>
__global__ void __launch_bounds__(256) kMagneticYee(FieldPtrsT<R> F, StepCoefsT<R> c, CpmlArgsT<R> P) {
  const int nx = F.n[0][0], ny = F.n[1][1], nz = F.n[2][2];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t % (nx + 1), j = t / (nx + 1);
  const int k = blockIdx.y;
  if (j > ny || k > nz) return;
  if (j < ny && k < nz) magneticBx<false>(F, c, P, i, j, k);
  if (i < nx && k < nz) magneticBy<false>(F, c, P, i, j, k);
  if (i < nx && j < ny)
    finishBz(F, c, P, i, j, k, rawE<1>(F, i + 1, j, k) - rawE<1>(F, i, j, k), rawE<0>(F, i, j + 1, k) - rawE<0>(F, i, j, k));
}

// Cowan edge nodes the tiles do not cover: Bx at i = nx and By at j = ny.
constexpr int kEdgeT = 32, kEdgeZ = 4;  // edge nodes per workgroup: 32 along the edge's free axis x 4 planes
template <class R>
__global__ void __launch_bounds__(256) kMagneticCowanEdges(FieldPtrsT<R> F, StepCoefsT<R> c, CpmlArgsT<R> P) {
  // Workgroups [0, nbX) take tiles of the Bx(i = nx) edge over (j, k), the
  // rest tiles of the By(j = ny) edge over (i, k). Each smoothed sample a
  // tile's curls read is computed once into LDS (the difference of two
  // neighbouring samples is each node's curl term), then each node is
  // finished from there: per tile, (T+1) x Z samples along the free axis
  // and T x (Z+1) along z.
  constexpr int T = kEdgeT, Z = kEdgeZ, NA = (T + 1) * Z, NB = T * (Z + 1);
  __shared__ double sa[NA], sb[NB];
  const int nx = F.n[0][0], ny = F.n[1][1], nz = F.n[2][2];
  const int gz = (nz + Z - 1) / Z;
  const int nbX = ((ny + T - 1) / T) * gz;
  const bool edgeX = static_cast<int>(blockIdx.x) < nbX;
  const int w = edgeX ? static_cast<int>(blockIdx.x) : static_cast<int>(blockIdx.x) - nbX;
  const int n = edgeX ? ny : nx;  // extent of the free axis (nodes 0..n-1)
  const int gu = (n + T - 1) / T;
  const int t0 = (w % gu) * T, k0 = (w / gu) * Z;
  for (int q = static_cast<int>(threadIdx.x); q < NA + NB; q += static_cast<int>(blockDim.x)) {
    if (q < NA) {
      // S_y Ez (x edge) / S_x Ez (y edge) at free index t0 + q % (T+1) <= n, plane k0 + q / (T+1) < nz
      const int u = t0 + q % (T + 1), k = k0 + q / (T + 1);
      if (u <= n && k < nz) sa[q] = edgeX ? smoothE<2, 1>(F, c, nx, u, k) : smoothE<2, 0>(F, c, u, ny, k);
    } else {
      // S_z Ey (x edge) / S_z Ex (y edge) at free index t0 + r % T < n, plane k0 + r / T <= nz
      const int r = q - NA;
      const int u = t0 + r % T, k = k0 + r / T;
      if (u < n && k <= nz) sb[r] = edgeX ? smoothE<1, 2>(F, c, nx, u, k) : smoothE<0, 2>(F, c, u, ny, k);
    }
  }
  __syncthreads();
  if (threadIdx.x >= T * Z) return;
  const int du = static_cast<int>(threadIdx.x) % T, dk = static_cast<int>(threadIdx.x) / T;
  const int u = t0 + du, k = k0 + dk;
  if (u >= n || k >= nz) return;
  const double across = sa[dk * (T + 1) + du + 1] - sa[dk * (T + 1) + du];  // dEz/dy (x edge), dEz/dx (y edge)
  const double along = sb[(dk + 1) * T + du] - sb[dk * T + du];              // dEy/dz (x edge), dEx/dz (y edge)
  if (edgeX)
    finishBx(F, c, P, nx, u, k, across, along);
  else
    finishBy(F, c, P, u, ny, k, along, across);
}

// Tile of the Cowan magnetic update: kMX x kMY nodes of a z plane, staged
// with a one-node halo (kLX x kLY) for each of three planes and three E
// components.
constexpr int kMX = 32, kMY = 8;
constexpr int kLX = kMX + 2, kLY = kMY + 2, kLP = kLX * kLY;
constexpr int kStageN = 9 * kLP;  // 3 components x 3 planes

// Smoothed sample from the staged planes; (li, lj) local.
// pl[0] = plane k-1, pl[1] = plane k, pl[2] = plane k+1; D is the derivative axis.
template <int D, class R>
__device__ __forceinline__ double smoothL(const R* const pl[3], const StepCoefsT<R>& c, int li, int lj) {
  constexpr int TA = (D + 1) % 3, TB = (D + 2) % 3;
  const double delta = D == 0 ? c.dYZ : (D == 1 ? c.dZX : c.dXY);
  auto at = [&](int dx, int dy, int dz) { return static_cast<double>(pl[1 + dz][(lj + dy) * kLX + li + dx]); };
  auto off = [&](int axis, int dir, int& dx, int& dy, int& dz) {
    if (axis == 0) dx += dir; else if (axis == 1) dy += dir; else dz += dir;
  };
  double r = c.alpha[D] * at(0, 0, 0);
  for (int dir = -1; dir <= 1; dir += 2) {
    int ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
    off(TA, dir, ax, ay, az);
    off(TB, dir, bx, by, bz);
    r += c.beta[TA] * at(ax, ay, az);
    r += c.beta[TB] * at(bx, by, bz);
  }
  for (int dA = -1; dA <= 1; dA += 2)
    for (int dB = -1; dB <= 1; dB += 2) {
      int x = 0, y = 0, z = 0;
      off(TA, dA, x, y, z);
      off(TB, dB, x, y, z);
      r += delta * at(x, y, z);
    }
  return r;
}

// evalE() split into its in-plane part and its z part, so that staging can
// compute addresses without loads and then issue every load at once.
// In-plane: the index j * n0 + i of the node evalE reads for (i, j) (PEC
// parity clamp on the component's own axis, odd reflection across the
// others) and its sign; sign 0 for nodes more than one node beyond the
// array, which feed no B node.
template <int A, class R>
__device__ __forceinline__ int planeIndexE(const FieldPtrsT<R>& F, int i, int j, int& sign) {
  const int n0 = F.n[A][0], n1 = F.n[A][1];
  sign = (i < -1 || i > n0 || j < -1 || j > n1) ? 0 : 1;
  if (i < 0) {
    if (A == 0) i = 0;
    else { i = -i; sign = -sign; }
  } else if (i >= n0) {
    if (A == 0) i = n0 - 1;
    else { i = 2 * (n0 - 1) - i; sign = -sign; }
  }
  if (j < 0) {
    if (A == 1) j = 0;
    else { j = -j; sign = -sign; }
  } else if (j >= n1) {
    if (A == 1) j = n1 - 1;
    else { j = 2 * (n1 - 1) - j; sign = -sign; }
  }
  return sign == 0 ? 0 : j * n0 + i;
}

// z part: the plane evalE reads for plane k (the array's own plane, the halo
// plane, or the PEC parity image) and its sign.
template <int A, class R>
__device__ __forceinline__ const R* planeBaseE(const FieldPtrsT<R>& F, int k, int& sign) {
  const int n0 = F.n[A][0], n1 = F.n[A][1], n2 = F.n[A][2];
  sign = 1;
  if (k >= 0 && k < n2) return F.e[A] + static_cast<size_t>(k) * n1 * n0;
  const bool lower = k < 0;
  const R* plane = lower ? F.lo[A] : F.up[A];
  if (plane) return plane;
  if (A == 2) k = lower ? 0 : n2 - 1;
  else { k = lower ? -k : 2 * (n2 - 1) - k; sign = -1; }
  return F.e[A] + static_cast<size_t>(k) * n1 * n0;
}

// z-marching Cowan magnetic update for i < nx, j < ny: one workgroup per
// (kMX x kMY) tile and run of kMZ z planes [kb, ke). A ring of three LDS
// planes per E component holds planes k-1, k, k+1 of the tile plus a
// one-node halo as electricValue() samples (reflection / halo / PEC parity
// applied while loading); plane k+2 is loaded into registers before plane
// k is computed and stored into the ring slot of plane k-1 afterwards, so
// every E plane of the run is read from memory once and its load latency
// overlaps the smoothing of the plane before. Each smoothed sample is the
// same nine products in the same order as smoothedElectric(), and each is
// computed once: a node's curl takes the difference of two samples at
// neighbouring nodes, so the x-z and y-z smoothed samples of the plane are
// formed once per node (plus one halo row or column) into LDS and the curl
// reads both from there, and the x-y smoothed sample of plane k+1 is kept
// in registers for the next plane. 6 samples per node instead of 12.
constexpr int kMZ = 8;
constexpr int kSX = kMX + 1, kSY = kMY + 1;  // sample grid: one halo column / row
template <class R>
__global__ void __launch_bounds__(kMX * kMY) kMagneticCowanMarch(FieldPtrsT<R> F, StepCoefsT<R> c, CpmlArgsT<R> P,
                                                                 int runZ) {
  __shared__ R buf[kStageN];  // [component][ring slot][kLY][kLX]
  // smoothed samples of the plane: [S_y Ez, S_y Ex, S_x Ez, S_x Ey][kSY][kSX]
  __shared__ double smp[4 * kSY * kSX];
  const int nx = F.n[0][0], ny = F.n[1][1], nz = F.n[2][2];
  // tiles x fastest, then y, then the z runs; an XCD takes a contiguous
  // share, so the halo rows and the runs' boundary planes its neighbours
  // read stay in its L2
  const int gx = (nx + kMX - 1) / kMX, gy = (ny + kMY - 1) / kMY;
  const int w = xcdContiguous(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
  const int i0 = (w % gx) * kMX, j0 = ((w / gx) % gy) * kMY;
  const int kb = (w / (gx * gy)) * runZ, ke = min(kb + runZ, nz + 1);
  const int tx = static_cast<int>(threadIdx.x) % kMX, ty = static_cast<int>(threadIdx.x) / kMX;
  const int i = i0 + tx, j = j0 + ty;
  // The region node (gi, gj) of a lane is the same on every plane: its
  // in-plane index and sign per component are computed once.
  constexpr int kRegionR = (kLP + kMX * kMY - 1) / (kMX * kMY);
  int ix[kRegionR][3], sxy[kRegionR][3];
#pragma unroll
  for (int r = 0; r < kRegionR; ++r) {
    const int q = static_cast<int>(threadIdx.x) + r * kMX * kMY;
    const int qc = q < kLP ? q : 0;
    const int gi = i0 - 1 + qc % kLX, gj = j0 - 1 + qc / kLX;
    ix[r][0] = planeIndexE<0>(F, gi, gj, sxy[r][0]);
    ix[r][1] = planeIndexE<1>(F, gi, gj, sxy[r][1]);
    ix[r][2] = planeIndexE<2>(F, gi, gj, sxy[r][2]);
  }
  // One plane's samples of this lane's region nodes, loads only.
  auto load = [&](int k, R v[kRegionR][3], int zs[3]) {
    const R* base[3];
    base[0] = planeBaseE<0>(F, k, zs[0]);
    base[1] = planeBaseE<1>(F, k, zs[1]);
    base[2] = planeBaseE<2>(F, k, zs[2]);
#pragma unroll
    for (int r = 0; r < kRegionR; ++r)
#pragma unroll
      for (int cp = 0; cp < 3; ++cp) v[r][cp] = base[cp][ix[r][cp]];
  };
  auto store = [&](int slot, const R v[kRegionR][3], const int zs[3]) {
#pragma unroll
    for (int r = 0; r < kRegionR; ++r) {
      const int q = static_cast<int>(threadIdx.x) + r * kMX * kMY;
      if (q < kLP)
#pragma unroll
        for (int cp = 0; cp < 3; ++cp) {
          const int sg = sxy[r][cp] * zs[cp];
          buf[(cp * 3 + slot) * kLP + q] = sg == 0 ? R(0) : (sg > 0 ? v[r][cp] : -v[r][cp]);
        }
    }
  };
  auto slotOf = [](int k) { return ((k % 3) + 3) % 3; };
  {
    R v0[kRegionR][3], v1[kRegionR][3], v2[kRegionR][3];
    int z0[3], z1[3], z2[3];
    load(kb - 1, v0, z0);
    load(kb, v1, z1);
    load(kb + 1, v2, z2);
    store(slotOf(kb - 1), v0, z0);
    store(slotOf(kb), v1, z1);
    store(slotOf(kb + 1), v2, z2);
  }
  __syncthreads();
  const int li = tx + 1, lj = ty + 1;
  const bool owner = i < nx && j < ny;
  double* const syEz = smp;
  double* const syEx = smp + kSY * kSX;
  double* const sxEz = smp + 2 * kSY * kSX;
  double* const sxEy = smp + 3 * kSY * kSX;
  const int at = ty * kSX + tx;
  // the x-y smoothed Ey, Ex of plane k + 1, kept for plane k + 1
  double zEy = 0.0, zEx = 0.0;
  bool zKept = false;
  // magnetic CPML ordinals: of the lane's i and j once, of each plane's k
  // one plane ahead
  const int mI = owner && P.enabled ? P.ord[0][0][i] : -1, mJ = owner && P.enabled ? P.ord[0][1][j] : -1;
  int mKn = owner && P.enabled ? P.ord[0][2][kb] : -1;
  for (int k = kb; k < ke; ++k) {
    // the plane's B loads (old values, CPML histories) first, so that the
    // update waits for them only and their latency hides behind the
    // smoothing; then plane k + 2's E loads
    const int mK = mKn;
    if (owner && P.enabled && k + 1 < ke) mKn = P.ord[0][2][k + 1];
    BLoads<R> lx{}, ly{}, lz{};
    if (owner) {
      if (k < nz) {
        lx = loadB<0>(F, P, i, j, k, -1, mJ, mK);
        ly = loadB<1>(F, P, i, j, k, mI, -1, mK);
      }
      lz = loadB<2>(F, P, i, j, k, mI, mJ, -1);
    }
    R next[kRegionR][3];
    int zn[3] = {0, 0, 0};
    const bool more = k + 1 < ke;  // plane k + 2 feeds plane k + 1 of this run
    if (more) load(k + 2, next, zn);
    const R* ex[3] = {buf + (0 * 3 + slotOf(k - 1)) * kLP, buf + (0 * 3 + slotOf(k)) * kLP,
                      buf + (0 * 3 + slotOf(k + 1)) * kLP};
    const R* ey[3] = {buf + (1 * 3 + slotOf(k - 1)) * kLP, buf + (1 * 3 + slotOf(k)) * kLP,
                      buf + (1 * 3 + slotOf(k + 1)) * kLP};
    const R* ez[3] = {buf + (2 * 3 + slotOf(k - 1)) * kLP, buf + (2 * 3 + slotOf(k)) * kLP,
                      buf + (2 * 3 + slotOf(k + 1)) * kLP};
    // Every lane's node of the tile, then the halo row (S_y samples at
    // j0 + kMY) and the halo column (S_x samples at i0 + kMX); all read the
    // staged region only.
    syEz[at] = smoothL<1>(ez, c, li, lj);
    syEx[at] = smoothL<1>(ex, c, li, lj);
    sxEz[at] = smoothL<0>(ez, c, li, lj);
    sxEy[at] = smoothL<0>(ey, c, li, lj);
    if (threadIdx.x < kMX) {
      const int h = kMY * kSX + tx;
      syEz[h] = smoothL<1>(ez, c, li, kMY + 1);
      syEx[h] = smoothL<1>(ex, c, li, kMY + 1);
    } else if (threadIdx.x < kMX + kMY) {
      const int r = static_cast<int>(threadIdx.x) - kMX;
      const int h = r * kSX + kMX;
      sxEz[h] = smoothL<0>(ez, c, kMX + 1, r + 1);
      sxEy[h] = smoothL<0>(ey, c, kMX + 1, r + 1);
    }
    __syncthreads();
    if (owner) {
      if (k < nz) {
        // S_z is x-y smoothing: only its own plane is read
        const R* eyk[3] = {ey[1], ey[1], ey[1]};
        const R* eyk1[3] = {ey[2], ey[2], ey[2]};
        const R* exk[3] = {ex[1], ex[1], ex[1]};
        const R* exk1[3] = {ex[2], ex[2], ex[2]};
        const double ey0 = zKept ? zEy : smoothL<2>(eyk, c, li, lj);
        const double ex0 = zKept ? zEx : smoothL<2>(exk, c, li, lj);
        zEy = smoothL<2>(eyk1, c, li, lj);
        zEx = smoothL<2>(exk1, c, li, lj);
        zKept = true;
        finishB<0>(F, c, P, i, j, k, syEz[at + kSX] - syEz[at], zEy - ey0, lx);
        finishB<1>(F, c, P, i, j, k, zEx - ex0, sxEz[at + 1] - sxEz[at], ly);
      }
      finishB<2>(F, c, P, i, j, k, sxEy[at + 1] - sxEy[at], syEx[at + kSX] - syEx[at], lz);
    }
    if (more) {
      __syncthreads();  // every lane is done with plane k - 1
      store(slotOf(k + 2), next, zn);
      __syncthreads();
    }
  }
}

template <class R>
__device__ __forceinline__ double rd(const R* p, int nx, int ny, int i, int j, int k) {
  return static_cast<double>(p[(static_cast<size_t>(k) * ny + j) * nx + i]);
}

// Fused electric update: EBFieldGrid::advanceElectric interior curl
// (eb_field.cpp:582-624), EBConvolutionalPML::afterElectricUpdate
// (eb_cpml.cpp:470-539) incl. its PEC termination (eb_field.cpp:158-204),
// and the moving-window inlet overwrite (moving_window_fields.cpp:276-280).
// MODE 0 writes E in place and advances the histories; MODE 1 writes the
// zero-current preview into `out` and leaves the histories untouched.
// E update at lattice node (i, j, k) of the (nx+1) x (ny+1) x (nz+1) node
// box: Ex (i < nx), Ey (j < ny) and Ez (k < nz) of that node. Every load of
// the three components (CPML ordinals, E, J, the curl's B samples, then the
// histories the ordinals select) is issued before the first store, so a
// thread has them in flight at once; the arithmetic is then done per
// component in the reference's order and the results stored.
template <class R>
__device__ __forceinline__ double cpmlTerm(const CpmlPointT<R>& c, double raw, R hOld, double& hNew) {
  hNew = c.decay * static_cast<double>(hOld) + c.drive * raw;  // transformedCorrection, eb_cpml.cpp:131-138
  return c.ikm1 * raw + hNew;
}

template <int MODE, class R>
__device__ __forceinline__ void electricNode(const FieldPtrsT<R>& F, const StepCoefsT<R>& c, const CpmlArgsT<R>& P,
                                             R* ox, R* oy, R* oz, R* sx, R* sy, R* sz, const R* inletX,
                                             const R* inletY, int i, int j, int k) {
  const int nx = F.n[0][0], ny = F.n[1][1], nz = F.n[2][2];
  const R* bx = F.b[0];  // (nx+1) x ny x nz
  const R* by = F.b[1];  // nx x (ny+1) x nz
  const R* bz = F.b[2];  // nx x ny x (nz+1)
  const int kb = P.lowerZ ? 1 : 0, ke = P.upperZ ? nz : nz + 1;
  const bool hasX = i < nx, hasY = j < ny, hasZ = k < nz;
  const size_t ix = (static_cast<size_t>(k) * (ny + 1) + j) * nx + i;        // Ex: nx x (ny+1) x (nz+1)
  const size_t iy = (static_cast<size_t>(k) * ny + j) * (nx + 1) + i;        // Ey: (nx+1) x ny x (nz+1)
  const size_t iz = (static_cast<size_t>(k) * (ny + 1) + j) * (nx + 1) + i;  // Ez: (nx+1) x (ny+1) x nz
  const bool curlX = hasX && k >= 1 && k < nz && j >= 1 && j < ny;
  const bool curlY = hasY && k >= 1 && k < nz && i >= 1 && i < nx;
  const bool curlZ = hasZ && i >= 1 && i < nx && j >= 1 && j < ny;
  const bool cpX = P.enabled && hasX && j >= 1 && j < ny && k >= kb && k < ke;
  const bool cpY = P.enabled && hasY && i >= 1 && i < nx && k >= kb && k < ke;
  const bool cpZ = P.enabled && hasZ && i >= 1 && i < nx && j >= 1 && j < ny;
  auto bxAt = [&](int a, int b2, int c2) { return bx[(static_cast<size_t>(c2) * ny + b2) * (nx + 1) + a]; };
  auto byAt = [&](int a, int b2, int c2) { return by[(static_cast<size_t>(c2) * (ny + 1) + b2) * nx + a]; };
  auto bzAt = [&](int a, int b2, int c2) { return bz[(static_cast<size_t>(c2) * ny + b2) * nx + a]; };

  // loads without dependencies
  int oI = -1, oJ = -1, oK = -1;  // CPML ordinals of i, j, k (-1: outside the layers)
  if (cpY || cpZ) oI = P.ord[1][0][i];
  if (cpX || cpZ) oJ = P.ord[1][1][j];
  if (cpX || cpY) oK = P.ord[1][2][k];
  R ex0 = 0, ey0 = 0, ez0 = 0, jx = 0, jy = 0, jz = 0;
  // curl samples: Ex bz(j), bz(j-1), by(k), by(k-1); Ey bx(k), bx(k-1),
  // bz(i), bz(i-1); Ez by(i), by(i-1), bx(j), bx(j-1)
  R xz0 = 0, xz1 = 0, xy0 = 0, xy1 = 0, yx0 = 0, yx1 = 0, yz0 = 0, yz1 = 0, zy0 = 0, zy1 = 0, zx0 = 0, zx1 = 0;
  if (hasX) ex0 = F.e[0][ix];
  if (hasY) ey0 = F.e[1][iy];
  if (hasZ) ez0 = F.e[2][iz];
  if (curlX) {
    xz0 = bzAt(i, j, k); xz1 = bzAt(i, j - 1, k); xy0 = byAt(i, j, k); xy1 = byAt(i, j, k - 1);
    jx = F.j[0][ix];
  }
  if (curlY) {
    yx0 = bxAt(i, j, k); yx1 = bxAt(i, j, k - 1); yz0 = bzAt(i, j, k); yz1 = bzAt(i - 1, j, k);
    jy = F.j[1][iy];
  }
  if (curlZ) {
    zy0 = byAt(i, j, k); zy1 = byAt(i - 1, j, k); zx0 = bxAt(i, j, k); zx1 = bxAt(i, j - 1, k);
    jz = F.j[2][iz];
  }
  R inX = 0, inY = 0;
  if (inletX && k == 0 && hasX) inX = inletX[static_cast<size_t>(j) * nx + i];
  if (inletY && k == 0 && hasY) inY = inletY[static_cast<size_t>(j) * (nx + 1) + i];

  // loads the CPML ordinals select: the samples a term needs outside the
  // curl's range and the histories
  const bool tX1 = cpX && oJ >= 0, tX2 = cpX && oK >= 0;
  const bool tY1 = cpY && oK >= 0, tY2 = cpY && oI >= 0;
  const bool tZ1 = cpZ && oI >= 0, tZ2 = cpZ && oJ >= 0;
  if (!curlX && tX1) { xz0 = bzAt(i, j, k); xz1 = bzAt(i, j - 1, k); }
  if (!curlX && tX2) { xy0 = byAt(i, j, k); xy1 = byAt(i, j, k - 1); }
  if (!curlY && tY1) { yx0 = bxAt(i, j, k); yx1 = bxAt(i, j, k - 1); }
  if (!curlY && tY2) { yz0 = bzAt(i, j, k); yz1 = bzAt(i - 1, j, k); }
  if (!curlZ && tZ1) { zy0 = byAt(i, j, k); zy1 = byAt(i - 1, j, k); }
  if (!curlZ && tZ2) { zx0 = bxAt(i, j, k); zx1 = bxAt(i, j - 1, k); }
  R* hp[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  R hv[6] = {0, 0, 0, 0, 0, 0};
  if (tX1) { hp[0] = histAt<1, 0, 1>(P, oJ, i, j, k, nx); hv[0] = *hp[0]; }
  if (tX2) { hp[1] = histAt<1, 0, 2>(P, oK, i, j, k, nx); hv[1] = *hp[1]; }
  if (tY1) { hp[2] = histAt<1, 1, 2>(P, oK, i, j, k, nx + 1); hv[2] = *hp[2]; }
  if (tY2) { hp[3] = histAt<1, 1, 0>(P, oI, i, j, k, nx + 1); hv[3] = *hp[3]; }
  if (tZ1) { hp[4] = histAt<1, 2, 0>(P, oI, i, j, k, nx + 1); hv[4] = *hp[4]; }
  if (tZ2) { hp[5] = histAt<1, 2, 1>(P, oJ, i, j, k, nx + 1); hv[5] = *hp[5]; }

  // arithmetic, EBFieldGrid::advanceElectric + afterElectricUpdate order
  double hn[6] = {0, 0, 0, 0, 0, 0};
  double ex = 0, ey = 0, ez = 0;
  if (hasX) {
    ex = static_cast<double>(ex0);
    const double dz = static_cast<double>(xz0) - static_cast<double>(xz1);
    const double dy = static_cast<double>(xy0) - static_cast<double>(xy1);
    if (curlX) ex += c.cY * dz - c.cZ * dy - c.cs * jx;
    if (tX1) ex += (1.0 * P.eleScale[1]) * cpmlTerm(P.prof[1][1][j], dz, hv[0], hn[0]);
    if (tX2) ex += (-1.0 * P.eleScale[2]) * cpmlTerm(P.prof[1][2][k], dy, hv[1], hn[1]);
    if (j == 0 || j == ny || (P.lowerZ && k == 0) || (P.upperZ && k == nz)) ex = 0.0;
    if (inletX && k == 0) ex = inX;
  }
  if (hasY) {
    ey = static_cast<double>(ey0);
    const double dx = static_cast<double>(yx0) - static_cast<double>(yx1);
    const double dz = static_cast<double>(yz0) - static_cast<double>(yz1);
    if (curlY) ey += c.cZ * dx - c.cX * dz - c.cs * jy;
    if (tY1) ey += (1.0 * P.eleScale[2]) * cpmlTerm(P.prof[1][2][k], dx, hv[2], hn[2]);
    if (tY2) ey += (-1.0 * P.eleScale[0]) * cpmlTerm(P.prof[1][0][i], dz, hv[3], hn[3]);
    if (i == 0 || i == nx || (P.lowerZ && k == 0) || (P.upperZ && k == nz)) ey = 0.0;
    if (inletY && k == 0) ey = inY;
  }
  if (hasZ) {
    ez = static_cast<double>(ez0);
    const double dy = static_cast<double>(zy0) - static_cast<double>(zy1);
    const double dx = static_cast<double>(zx0) - static_cast<double>(zx1);
    if (curlZ) ez += c.cX * dy - c.cY * dx - c.cs * jz;
    if (tZ1) ez += (1.0 * P.eleScale[0]) * cpmlTerm(P.prof[1][0][i], dy, hv[4], hn[4]);
    if (tZ2) ez += (-1.0 * P.eleScale[1]) * cpmlTerm(P.prof[1][1][j], dx, hv[5], hn[5]);
    if (i == 0 || i == nx || j == 0 || j == ny) ez = 0.0;
  }

  // stores (MODE 1 keeps the histories)
  if (MODE == 0) {
#pragma unroll
    for (int t = 0; t < 6; ++t)
      if (hp[t]) *hp[t] = static_cast<R>(hn[t]);
  }
  if (hasX) {
    (MODE == 0 ? F.e[0] : ox)[ix] = static_cast<R>(ex);
    if (MODE == 1 && sx) sx[ix] = static_cast<R>(ex + static_cast<double>(ex0));
  }
  if (hasY) {
    (MODE == 0 ? F.e[1] : oy)[iy] = static_cast<R>(ey);
    if (MODE == 1 && sy) sy[iy] = static_cast<R>(ey + static_cast<double>(ey0));
  }
  if (hasZ) {
    (MODE == 0 ? F.e[2] : oz)[iz] = static_cast<R>(ez);
    if (MODE == 1 && sz) sz[iz] = static_cast<R>(ez + static_cast<double>(ez0));
  }
}

// The same update component after component, each stored before the next
// one's loads. The whole-grid E update in FP32 storage keeps this order: its
// loads are half as wide and its 27 registers allow seven waves per SIMD,
// where the two-phase form needs 90 registers and runs slower (43 against
// 39 us on a 64x64x512 grid; 59 against 64 us in FP64).
template <int MODE, class R>
__device__ __forceinline__ void electricNodeOnePass(const FieldPtrsT<R>& F, const StepCoefsT<R>& c, const CpmlArgsT<R>& P,
                                             R* ox, R* oy, R* oz, R* sx, R* sy, R* sz, const R* inletX,
                                             const R* inletY, int i, int j, int k) {
  const int nx = F.n[0][0], ny = F.n[1][1], nz = F.n[2][2];
  constexpr bool W = MODE == 0;
  const R* bx = F.b[0];
  const R* by = F.b[1];
  const R* bz = F.b[2];
  const int kb = P.lowerZ ? 1 : 0, ke = P.upperZ ? nz : nz + 1;
  if (i < nx) {  // Ex: nx x (ny+1) x (nz+1)
    const size_t idx = (static_cast<size_t>(k) * (ny + 1) + j) * nx + i;
    const double e0 = static_cast<double>(F.e[0][idx]);
    double e = e0;
    if (k >= 1 && k < nz && j >= 1 && j < ny)
      e += c.cY * (rd(bz, nx, ny, i, j, k) - rd(bz, nx, ny, i, j - 1, k)) -
           c.cZ * (rd(by, nx, ny + 1, i, j, k) - rd(by, nx, ny + 1, i, j, k - 1)) - c.cs * F.j[0][idx];
    if (P.enabled && j >= 1 && j < ny && k >= kb && k < ke) {
      int o = P.ord[1][1][j];
      if (o >= 0)
        e += (1.0 * P.eleScale[1]) * cpmlCorrection<W>(P.prof[1][1][j], rd(bz, nx, ny, i, j, k) - rd(bz, nx, ny, i, j - 1, k),
                                                        histAt<1, 0, 1>(P, o, i, j, k, nx));
      o = P.ord[1][2][k];
      if (o >= 0)
        e += (-1.0 * P.eleScale[2]) * cpmlCorrection<W>(P.prof[1][2][k], rd(by, nx, ny + 1, i, j, k) - rd(by, nx, ny + 1, i, j, k - 1),
                                                         histAt<1, 0, 2>(P, o, i, j, k, nx));
    }
    if (j == 0 || j == ny || (P.lowerZ && k == 0) || (P.upperZ && k == nz)) e = 0.0;
    if (inletX && k == 0) e = inletX[static_cast<size_t>(j) * nx + i];
    (MODE == 0 ? F.e[0] : ox)[idx] = static_cast<R>(e);
    if (MODE == 1 && sx) sx[idx] = static_cast<R>(e + e0);
  }
  if (j < ny) {  // Ey: (nx+1) x ny x (nz+1)
    const size_t idx = (static_cast<size_t>(k) * ny + j) * (nx + 1) + i;
    const double e0 = static_cast<double>(F.e[1][idx]);
    double e = e0;
    if (k >= 1 && k < nz && i >= 1 && i < nx)
      e += c.cZ * (rd(bx, nx + 1, ny, i, j, k) - rd(bx, nx + 1, ny, i, j, k - 1)) -
           c.cX * (rd(bz, nx, ny, i, j, k) - rd(bz, nx, ny, i - 1, j, k)) - c.cs * F.j[1][idx];
    if (P.enabled && i >= 1 && i < nx && k >= kb && k < ke) {
      int o = P.ord[1][2][k];
      if (o >= 0)
        e += (1.0 * P.eleScale[2]) * cpmlCorrection<W>(P.prof[1][2][k], rd(bx, nx + 1, ny, i, j, k) - rd(bx, nx + 1, ny, i, j, k - 1),
                                                        histAt<1, 1, 2>(P, o, i, j, k, nx + 1));
      o = P.ord[1][0][i];
      if (o >= 0)
        e += (-1.0 * P.eleScale[0]) * cpmlCorrection<W>(P.prof[1][0][i], rd(bz, nx, ny, i, j, k) - rd(bz, nx, ny, i - 1, j, k),
                                                         histAt<1, 1, 0>(P, o, i, j, k, nx + 1));
    }
    if (i == 0 || i == nx || (P.lowerZ && k == 0) || (P.upperZ && k == nz)) e = 0.0;
    if (inletY && k == 0) e = inletY[static_cast<size_t>(j) * (nx + 1) + i];
    (MODE == 0 ? F.e[1] : oy)[idx] = static_cast<R>(e);
    if (MODE == 1 && sy) sy[idx] = static_cast<R>(e + e0);
  }
  if (k < nz) {  // Ez: (nx+1) x (ny+1) x nz
    const size_t idx = (static_cast<size_t>(k) * (ny + 1) + j) * (nx + 1) + i;
    const double e0 = static_cast<double>(F.e[2][idx]);
    double e = e0;
    const bool in = i >= 1 && i < nx && j >= 1 && j < ny;
    if (in)
      e += c.cX * (rd(by, nx, ny + 1, i, j, k) - rd(by, nx, ny + 1, i - 1, j, k)) -
           c.cY * (rd(bx, nx + 1, ny, i, j, k) - rd(bx, nx + 1, ny, i, j - 1, k)) - c.cs * F.j[2][idx];
    if (P.enabled && in) {
      int o = P.ord[1][0][i];
      if (o >= 0)
        e += (1.0 * P.eleScale[0]) * cpmlCorrection<W>(P.prof[1][0][i], rd(by, nx, ny + 1, i, j, k) - rd(by, nx, ny + 1, i - 1, j, k),
                                                        histAt<1, 2, 0>(P, o, i, j, k, nx + 1));
      o = P.ord[1][1][j];
      if (o >= 0)
        e += (-1.0 * P.eleScale[1]) * cpmlCorrection<W>(P.prof[1][1][j], rd(bx, nx + 1, ny, i, j, k) - rd(bx, nx + 1, ny, i, j - 1, k),
                                                         histAt<1, 2, 1>(P, o, i, j, k, nx + 1));
    }
    if (i == 0 || i == nx || j == 0 || j == ny) e = 0.0;
    (MODE == 0 ? F.e[2] : oz)[idx] = static_cast<R>(e);
    if (MODE == 1 && sz) sz[idx] = static_cast<R>(e + e0);
  }
}

template <int MODE, class R>
__global__ void __launch_bounds__(256) kElectric(FieldPtrsT<R> F, StepCoefsT<R> c, CpmlArgsT<R> P, R* ox, R* oy,
                                                  R* oz, R* sx, R* sy, R* sz, const R* inletX,
                                                  const R* inletY, const unsigned* gate) {
  if (gate) {
    const unsigned g = *gate;
    if (g == 0u || (g >> 31)) return;
  }
  // one thread per node of the (nx+1) x (ny+1) plane k; the blocks of a
  // plane and of the planes next to it run on one XCD, whose L2 then holds
  // the B plane k - 1 the curl reads
  const int nx = F.n[0][0], ny = F.n[1][1], nz = F.n[2][2];
  const int perPlane = (((nx + 1) * (ny + 1)) + blockDim.x - 1) / blockDim.x;
  const int w = xcdContiguous(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
  const int t = (w % perPlane) * blockDim.x + threadIdx.x;
  const int i = t % (nx + 1), j = t / (nx + 1);
  const int k = w / perPlane;
  if (j > ny || k > nz) return;
  if constexpr (sizeof(R) == 8)
    electricNode<MODE>(F, c, P, ox, oy, oz, sx, sy, sz, inletX, inletY, i, j, k);
  else
    electricNodeOnePass<MODE>(F, c, P, ox, oy, oz, sx, sy, sz, inletX, inletY, i, j, k);
}

// The zero-current preview (MODE 1 of kElectric) on the nodes owned by the
// listed cell tiles only: count = the active field of *total, the tiles in list
// order, a workgroup per tile (looping when the grid is smaller). Every
// node value is kElectric<1>'s.
template <class R>
__global__ void __launch_bounds__(256) kElectricPreviewTiles(FieldPtrsT<R> F, StepCoefsT<R> c, CpmlArgsT<R> P,
                                                             R* ox, R* oy, R* oz, R* sx, R* sy, R* sz,
                                                             const R* inletX, const R* inletY, const int* tiles,
                                                             const unsigned long long* total, TileGrid tg) {
  const int nx = F.n[0][0], ny = F.n[1][1], nz = F.n[2][2];
  const int count = activeField(*total);
  for (int b = blockIdx.x; b < count; b += gridDim.x) {
    // neighbouring tiles of the list on one XCD (when one pass covers it)
    const int tile = tiles[count <= static_cast<int>(gridDim.x) ? xcdContiguous(b, count) : b];
    const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
    const int i0 = tx * tg.tx, i1 = tx == tg.ntx - 1 ? nx + 1 : min(nx + 1, i0 + tg.tx);
    const int j0 = ty * tg.ty, j1 = ty == tg.nty - 1 ? ny + 1 : min(ny + 1, j0 + tg.ty);
    const int k0 = tz * tg.tz, k1 = tz == tg.ntz - 1 ? nz + 1 : min(nz + 1, k0 + tg.tz);
    const int wx = i1 - i0, wy = j1 - j0, wz = k1 - k0;
    for (int t = threadIdx.x; t < wx * wy * wz; t += blockDim.x) {
      const int i = i0 + t % wx, j = j0 + (t / wx) % wy, k = k0 + t / (wx * wy);
      electricNode<1>(F, c, P, ox, oy, oz, sx, sy, sz, inletX, inletY, i, j, k);
    }
  }
}

template <class R>
__global__ void kFill(R* p, size_t n, R v) {
  for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[t] = v;
}

template <class R>
static void fill(hipStream_t s, R* p, size_t n, R v) {
  if (!n) return;
  const unsigned blocks = std::min<size_t>((n + 255) / 256, 4096);
  kFill<R><<<blocks, 256, 0, s>>>(p, n, v);
  launchCheck("fill");
}

template <class R>
void launchMagnetic(FieldStateT<R>& s, const FieldPtrsT<R>& f) {
  const StepCoefsT<R> c = s.coefs();
  const CpmlArgsT<R> P = s.cpmlArgs();
  ScopedKernel sk(s.profiler, s.stream, "magnetic");
  if (s.g.solver == 1) {
    // z runs as long as lets every workgroup be resident in one round (a
    // second partial round would double the sweep's time); at least kMZ
    // planes per run, whose ends re-read two planes each
    const int tiles = static_cast<int>(ceilDiv(s.g.nx, kMX) * ceilDiv(s.g.ny, kMY));
    const int planes = static_cast<int>(s.g.nz + 1);
    const int slots = residentBlocks(reinterpret_cast<const void*>(&kMagneticCowanMarch<R>), kMX * kMY);
    const int runs = std::max(1, std::min(planes, slots / tiles));
    const int oneRound = std::max(kMZ, static_cast<int>(ceilDiv(planes, runs)));
    // the one-round length, and longer and shorter runs, are tuned per grid
    std::vector<int> lengths{oneRound};
    for (int z : {kMZ, 2 * kMZ, 4 * kMZ, 2 * oneRound})
      if (z <= planes && std::find(lengths.begin(), lengths.end(), z) == lengths.end()) lengths.push_back(z);
    TuneScope tune(shapeKey("cowan_march", s.g, sizeof(R)), lengths, s.stream);
    const int runZ = tune.value();
    dim3 tg(static_cast<unsigned>(tiles) * ceilDiv(planes, runZ), 1, 1);
    kMagneticCowanMarch<R><<<tg, kMX * kMY, 0, s.stream>>>(f, c, P, runZ);
    kMagneticCowanEdges<R><<<(ceilDiv(s.g.ny, kEdgeT) + ceilDiv(s.g.nx, kEdgeT)) * ceilDiv(s.g.nz, kEdgeZ), 256, 0,
                              s.stream>>>(f, c, P);
  } else {
    dim3 grid(ceilDiv((s.g.nx + 1) * (s.g.ny + 1), 256), static_cast<unsigned>(s.g.nz + 1));
    kMagneticYee<R><<<grid, 256, 0, s.stream>>>(f, c, P);
  }
  launchCheck("magnetic");
}

template <class R>
void launchElectric(FieldStateT<R>& s, const FieldPtrsT<R>& f, int mode, R* const out[3], const R* inletX,
                    const R* inletY, R* const sum[3], const unsigned* gate) {
  const StepCoefsT<R> c = s.coefs();
  CpmlArgsT<R> P = s.cpmlArgs();
  if (!s.hasCpml) {  // PerfectElectricConductorBoundary(lowerZ, upperZ)
    P.lowerZ = s.lowerZ;
    P.upperZ = s.upperZ;
  }
  // workgroup size tuned per grid (one thread per node either way)
  // (a gated launch may skip its work, so it is never timed)
  TuneScope tune(shapeKey(mode == 0 ? "electric" : "electric_preview_grid", s.g, sizeof(R)), {256, 128, 64},
                 s.stream, gate == nullptr);
  dim3 block(static_cast<unsigned>(tune.value()), 1, 1);
  dim3 grid(ceilDiv((s.g.nx + 1) * (s.g.ny + 1), block.x) * static_cast<unsigned>(s.g.nz + 1), 1, 1);
  if (mode == 0) {
    ScopedKernel sk(s.profiler, s.stream, "electric");
    kElectric<0, R><<<grid, block, 0, s.stream>>>(f, c, P, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, inletX,
                                               inletY, gate);
  } else {
    ScopedKernel sk(s.profiler, s.stream, "electric_preview");
    kElectric<1, R><<<grid, block, 0, s.stream>>>(f, c, P, out[0], out[1], out[2], sum ? sum[0] : nullptr,
                                               sum ? sum[1] : nullptr, sum ? sum[2] : nullptr, inletX, inletY, gate);
  }
  launchCheck("electric");
}

template <class R>
void launchElectricPreviewTiles(FieldStateT<R>& s, const FieldPtrsT<R>& f, R* const out[3], R* const sum[3],
                                const R* inletX, const R* inletY, const int* tiles, const unsigned long long* total,
                                const TileGrid& tg) {
  const StepCoefsT<R> c = s.coefs();
  CpmlArgsT<R> P = s.cpmlArgs();
  if (!s.hasCpml) {
    P.lowerZ = s.lowerZ;
    P.upperZ = s.upperZ;
  }
  const unsigned blocks = static_cast<unsigned>(std::min(tg.ntx * tg.nty * tg.ntz, 4096));
  ScopedKernel sk(s.profiler, s.stream, "electric_preview");
  kElectricPreviewTiles<R><<<std::max(1u, blocks), 256, 0, s.stream>>>(f, c, P, out[0], out[1], out[2], sum[0],
                                                                        sum[1], sum[2], inletX, inletY, tiles,
                                                                        total, tg);
  launchCheck("electric preview tiles");
}

// Zero the owned J nodes of every listed cell tile (the tiles with an
// occupied neighbour tile: the only nodes a tile deposition can write).
template <class R>
__global__ void __launch_bounds__(256) kClearCurrentTiles(R* jx, R* jy, R* jz, FieldPtrsT<R> F,
                                                          CurrentFootprint fp) {
  const int tile = fp.active[blockIdx.x];
  const int tx = tile % fp.ntx, ty = (tile / fp.ntx) % fp.nty, tz = tile / (fp.ntx * fp.nty);
  for (int d = 0; d < 3; ++d) {
    const int n0 = F.n[6 + d][0], n1 = F.n[6 + d][1], n2 = F.n[6 + d][2];
    const int i0 = tx * fp.tx, i1 = tx == fp.ntx - 1 ? n0 : min(n0, i0 + fp.tx);
    const int j0 = ty * fp.ty, j1 = ty == fp.nty - 1 ? n1 : min(n1, j0 + fp.ty);
    const int k0 = tz * fp.tz, k1 = tz == fp.ntz - 1 ? n2 : min(n2, k0 + fp.tz);
    const int wx = i1 - i0, wy = j1 - j0, wz = k1 - k0;
    R* dst = d == 0 ? jx : (d == 1 ? jy : jz);
    for (int t = threadIdx.x; t < wx * wy * wz; t += blockDim.x) {
      const int i = i0 + t % wx, j = j0 + (t / wx) % wy, k = k0 + t / (wx * wy);
      dst[(static_cast<size_t>(k) * n1 + j) * n0 + i] = 0.0;
    }
  }
}

template <class R>
static void clearCurrentNow(FieldStateT<R>& s) {
  if (s.footprint.valid) {
    const CurrentFootprint& fp = s.footprint;
    if (fp.count > 0) {
      kClearCurrentTiles<R><<<fp.count, 256, 0, s.stream>>>(s.base(6), s.base(7), s.base(8), s.ptrs(), fp);
      launchCheck("clear current tiles");
    }
  } else {
    for (int a = 6; a < 9; ++a) fill(s.stream, s.comp[a].get(), s.compSize(a), R(0));
  }
  s.currentZero = true;
  s.footprint.valid = false;
}

template <class R>
void launchClearCurrent(FieldStateT<R>& s) {
  if (s.currentZero) return;  // J already holds zeros
  ScopedKernel sk(s.profiler, s.stream, "clear_current");
  clearCurrentNow(s);
}

// Recenter the slid arrays into their slack once the slack is used up.
template <class R>
static void recenter(FieldStateT<R>& s) {
  for (int a = 0; a < 6; ++a) {
    const size_t bytes = s.compSize(a) * sizeof(R);
    hipCheck(hipMemcpyAsync(s.scratch.get(), s.base(a), bytes, hipMemcpyDeviceToDevice, s.stream), "hipMemcpyAsync");
    hipCheck(hipMemcpyAsync(s.comp[a].get() + s.slack * s.plane(a), s.scratch.get(), bytes, hipMemcpyDeviceToDevice,
                            s.stream),
             "hipMemcpyAsync");
  }
  for (int h = 0; h < 12; ++h) {
    if (!s.histSlides[h] || !s.hist[h].size()) continue;
    const size_t block = s.histLayers[h] * s.histRow[h];
    for (size_t slot = 0; slot < s.histSlots[h]; ++slot) {
      R* slotBase = s.hist[h].get() + slot * (s.histLayers[h] + s.slack) * s.histRow[h];
      hipCheck(hipMemcpyAsync(s.scratch.get(), slotBase + s.zb * s.histRow[h], block * sizeof(R),
                              hipMemcpyDeviceToDevice, s.stream),
               "hipMemcpyAsync");
      hipCheck(hipMemcpyAsync(slotBase + s.slack * s.histRow[h], s.scratch.get(), block * sizeof(R),
                              hipMemcpyDeviceToDevice, s.stream),
               "hipMemcpyAsync");
    }
  }
  s.zb = s.slack;
}

template <class R>
struct ShiftArgs {
  R* dst[6];
  const R* src[6];
  int planeLen[6];
  R* hist[12];
  long long slotStride[12];
  int row[12], slots[12];
};

// Insert the new k = 0 planes of the six E/B components and zero the new
// layer of every transverse CPML history, in one launch.
template <class R>
__global__ void kShiftInsert(ShiftArgs<R> A) {
  const int item = blockIdx.y;  // 0..5 component planes, 6.. history slots
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (item < 6) {
    if (t < A.planeLen[item]) A.dst[item][t] = A.src[item] ? A.src[item][t] : 0.0;
    return;
  }
  int rest = item - 6;
  for (int h = 0; h < 12; ++h) {
    if (rest < A.slots[h]) {
      if (t < A.row[h]) A.hist[h][rest * A.slotStride[h] + t] = 0.0;
      return;
    }
    rest -= A.slots[h];
  }
}

// MovingWindowFields::Impl::shift (moving_window_fields.cpp:439-518) with
// resetCurrent: slide every E/B component and transverse history one plane,
// insert the reference planes (zero histories) at k = 0, clear J.
template <class R>
void windowShift(FieldStateT<R>& s, const R* const planes[6]) {
  ScopedKernel sk(s.profiler, s.stream, "shift");
  if (s.zb == 0) recenter(s);
  s.zb -= 1;
  ShiftArgs<R> A{};
  int maxLen = 0, items = 6;
  for (int a = 0; a < 6; ++a) {
    A.dst[a] = s.base(a);
    A.src[a] = planes ? planes[a] : nullptr;
    A.planeLen[a] = static_cast<int>(s.plane(a));
    maxLen = std::max(maxLen, A.planeLen[a]);
  }
  for (int h = 0; h < 12; ++h) {
    const bool on = s.histSlides[h] && s.hist[h].size();
    A.slots[h] = on ? static_cast<int>(s.histSlots[h]) : 0;
    A.row[h] = static_cast<int>(s.histRow[h]);
    A.hist[h] = on ? s.hist[h].get() + s.historyDeviceOffset(h, 0) : nullptr;
    A.slotStride[h] = static_cast<long long>((s.histLayers[h] + s.slack) * s.histRow[h]);
    items += A.slots[h];
    maxLen = std::max(maxLen, A.row[h]);
  }
  dim3 grid(ceilDiv(maxLen, 256), static_cast<unsigned>(items));
  kShiftInsert<R><<<grid, 256, 0, s.stream>>>(A);
  launchCheck("shift insert");
  if (!s.currentZero) clearCurrentNow(s);
}

#define APRL_FIELDS_INSTANTIATE(R)                                                                            \
  template void allocateFields<R>(FieldStateT<R>&, std::size_t);                                              \
  template void installCpml<R>(FieldStateT<R>&, const aprl_cpml_parameters&, std::size_t, std::size_t, bool, \
                               bool);                                                                         \
  template struct FieldStateT<R>;                                                                            \
  template void launchMagnetic<R>(FieldStateT<R>&, const FieldPtrsT<R>&);                                     \
  template void launchElectric<R>(FieldStateT<R>&, const FieldPtrsT<R>&, int, R* const*, const R*, const R*,  \
                                  R* const*, const unsigned*);                                                \
  template void launchClearCurrent<R>(FieldStateT<R>&);                                                       \
  template void launchElectricPreviewTiles<R>(FieldStateT<R>&, const FieldPtrsT<R>&, R* const*, R* const*,     \
                                              const R*, const R*, const int*, const unsigned long long*,       \
                                              const TileGrid&);                                                \
  template void windowShift<R>(FieldStateT<R>&, const R* const*);

}  // namespace aprl_gpu
