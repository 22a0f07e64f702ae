// Source-shape machinery shared by the FP64 and FP32-split particle paths:
// cell-tile node boxes, the chord crossing cuts, the binomial3-x and
// compensated-quintic weights and the tap visitor of one chord segment.
//
// Every function is a template on the real type R. With R = double the
// operations and their association order are those of the reference
// (longitudinal_shape.h), so the FP64 path stays bitwise; the FP32-split
// path instantiates them with R = float on coordinates taken relative to
// the particle's own start cell.
#pragma once

#include "tilecounts.hpp"

#include "particles.hpp"

namespace aprl_gpu {

enum : int { kShapeQuintic = 0, kShapeBinomial3X = 2 };

// Lanes per particle in the deposits: each lane of a particle visits one
// slice of its chord's taps (segmentTaps), the lanes of one slice forming
// whole waves. The quintic's 368 taps per segment are split over its four
// Gauss points; binomial3-x's 64 taps cost less than the chord set-up each
// lane would repeat, so it keeps one lane. The fixed-point LDS sums do not
// depend on which lane adds a tap.
template <int SHAPE>
constexpr int depositLanes() {
  return SHAPE == kShapeQuintic ? 4 : 1;
}

// Extended node box of a cell tile per shape: every tap of a chord that
// starts in the tile and moves less than one cell per axis lies inside.
// (X0, Y0, Z0) is the box origin below the tile origin; RZ is how many
// neighbour tiles the box reaches along z (one along x and y).
template <int SHAPE>
struct Box;
template <>
struct Box<kShapeBinomial3X> {
  // binomialXZSegment taps: x [i-1, i+2], y [j, j+1], z [k-1, k+2]
  static constexpr int X0 = 2, Y0 = 1, Z0 = 2;
  static constexpr int NX = kTileX + 5, NY = kTileY + 3, NZ = kTileZ + 5;
  static constexpr int RZ = 1;
  // each chord's weights sum to at most 2 in absolute value per component
  static constexpr double kWeightBound = 2.0;
};
template <>
struct Box<kShapeQuintic> {
  // quintic Gauss-point taps: x [i, i+1], y [j, j+1], z [k-5, k+6]
  static constexpr int X0 = 1, Y0 = 1, Z0 = 6;
  static constexpr int NX = kTileX + 3, NY = kTileY + 3, NZ = kTileZ + 13;
  static constexpr int RZ = 2;
  // the compensated quintic's coefficients sum to about 3.01 in absolute value
  static constexpr double kWeightBound = 4.0;
};
template <int SHAPE>
constexpr int boxNodes() {
  return Box<SHAPE>::NX * Box<SHAPE>::NY * Box<SHAPE>::NZ;
}

template <int SHAPE>
struct TileBox {
  int x0, y0, z0;
  __device__ int local(int i, int j, int k, unsigned& err) const {
    const unsigned li = static_cast<unsigned>(i - x0), lj = static_cast<unsigned>(j - y0),
                   lk = static_cast<unsigned>(k - z0);
    if (li >= static_cast<unsigned>(Box<SHAPE>::NX) || lj >= static_cast<unsigned>(Box<SHAPE>::NY) ||
        lk >= static_cast<unsigned>(Box<SHAPE>::NZ)) {
      err |= kErrTileHalo;
      return -1;
    }
    return static_cast<int>((lk * Box<SHAPE>::NY + lj) * Box<SHAPE>::NX + li);
  }
  // box contains [i, i + wx) x [j, j + wy) x [k, k + wz)
  __device__ bool holds(long long i, long long j, long long k, int wx, int wy, int wz) const {
    return i >= x0 && i + wx <= x0 + Box<SHAPE>::NX && j >= y0 && j + wy <= y0 + Box<SHAPE>::NY && k >= z0 &&
           k + wz <= z0 + Box<SHAPE>::NZ;
  }
  __device__ int base(long long i, long long j, long long k) const {
    return (static_cast<int>(k - z0) * Box<SHAPE>::NY + static_cast<int>(j - y0)) * Box<SHAPE>::NX +
           static_cast<int>(i - x0);
  }
};
template <int SHAPE>
__device__ __forceinline__ TileBox<SHAPE> tileBox(int tile, int ntx, int nty) {
  const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
  return TileBox<SHAPE>{tx * kTileX - Box<SHAPE>::X0, ty * kTileY - Box<SHAPE>::Y0, tz * kTileZ - Box<SHAPE>::Z0};
}

// Binomial3 node weights at in-cell offset f (source_coupling_math.h:105-112).
template <class R>
__device__ __forceinline__ void nodeWeights(R f, R w[4]) {
  w[0] = R(0.25) * (R(1.) - f);
  w[1] = R(0.5) - R(0.25) * f;
  w[2] = R(0.25) + R(0.25) * f;
  w[3] = R(0.25) * f;
}

// Crossing cuts of longitudinal_shape::weights (longitudinal_shape.h:100-103):
// chord start x and displacement v in lattice units; the face crossings as
// chord fractions, with 0 and 1, in ascending order.
constexpr int kMaxCuts = 10;
template <class R>
__device__ int orbitCuts(const R x[3], const R v[3], R cuts[kMaxCuts], unsigned& err) {
  int count = 2;
  cuts[0] = R(0.);
  cuts[1] = R(1.);
  for (int a = 0; a < 3; ++a) {
    if (v[a] != R(0.)) {
      const R lo = fmin(x[a], x[a] + v[a]), hi = fmax(x[a], x[a] + v[a]);
      for (long long j = static_cast<long long>(floor(lo)) + 1; j < hi; ++j) {
        if (count == kMaxCuts) { err |= kErrCrossings; return count; }
        cuts[count++] = (static_cast<R>(j) - x[a]) / v[a];
      }
    }
  }
  for (int s = 1; s < count; ++s) {  // insertion sort, same order as std::sort
    const R key = cuts[s];
    int t = s - 1;
    while (t >= 0 && cuts[t] > key) { cuts[t + 1] = cuts[t]; --t; }
    cuts[t + 1] = key;
  }
  return count;
}

// Gauss-Legendre nodes and weights of longitudinal_shape::weights
// (longitudinal_shape.h:105-106).
__device__ __forceinline__ double gaussNode(int t) {
  return t == 0 ? -.8611363115940525752 : (t == 1 ? -.3399810435848562648 : (t == 2 ? .3399810435848562648 : .8611363115940525752));
}
__device__ __forceinline__ double gaussWeight(int t) {
  return (t == 0 || t == 3) ? .3478548451374538574 : .6521451548625461426;
}

// longitudinal_shape::basis for the compensated quintic, longitudinal_shape.h:
// 25-55: the degree-5 (nodes) or degree-4 (edges) B-spline at offset f,
// convolved with the seven-point compensation.
template <class R>
struct QBasisT {
  int first, count;
  R v[12];
};
template <class R>
__device__ __forceinline__ void quinticBasis(R z, bool edge, QBasisT<R>& r) {
#pragma unroll
  for (int k = 0; k < 12; ++k) r.v[k] = R(0.0);
  const unsigned degree = edge ? 4u : 5u;
  const R cell = floor(z), f = z - cell;
  R n[6] = {R(1.), R(0.), R(0.), R(0.), R(0.), R(0.)};
#pragma unroll
  for (unsigned j = 1; j <= 5; ++j) {
    if (j > degree) break;
    R saved = R(0.);
#pragma unroll
    for (unsigned k = 0; k < j; ++k) {
      const R t = n[k] / static_cast<R>(j);
      n[k] = saved + (static_cast<R>(k + 1) - f) * t;
      saved = (f + static_cast<R>(j) - static_cast<R>(k) - R(1.)) * t;
    }
    n[j] = saved;
  }
  const double p = 5., a = p / 6., b = 3 * p / 40. + p * (p - 1) / 72.,
               d = 5 * p / 112. + p * (p - 1) / 80. + p * (p - 1) * (p - 2) / 1296.;
  const R c[7] = {R(-d / 64.), R(b / 16. + 3 * d / 32.), R(-a / 4. - b / 4. - 15 * d / 64.),
                  R(1 + a / 2. + 3 * b / 8. + 5 * d / 16.), R(-a / 4. - b / 4. - 15 * d / 64.),
                  R(b / 16. + 3 * d / 32.), R(-d / 64.)};
  r.first = static_cast<int>(cell) - 5;
  r.count = static_cast<int>(degree) + 7;
#pragma unroll
  for (unsigned j = 0; j <= 5; ++j) {
    if (j > degree) break;
#pragma unroll
    for (unsigned k = 0; k < 7; ++k) r.v[j + k] += n[j] * c[k];
  }
}

// Visit every tap of one chord segment [first, last] of
// longitudinal_shape::weights (longitudinal_shape.h:97-117) in the
// reference's order: binomialXZSegment (61-94) for binomial3-x, four Gauss
// points with the quintic basis otherwise. The chord starts at x and moves
// by v, in lattice units relative to the node org (grid index of x's
// origin); taps are reported with grid indices: visit(d, i, j, k, w). fast
// is called instead, with offsets from the segment's lowest tap corner,
// when the caller reports (through interior) that every tap of the segment
// lies inside its domain.
//
// With SPLIT > 1 only the taps of slice `sub` (0 <= sub < SPLIT) are
// visited, so SPLIT lanes can share one segment: binomial3-x slices the
// Jx / Jy taps by z offset c (c % SPLIT == sub) and the Jz taps by their
// (y, x) pair ((4 b + a) % SPLIT == sub), 16 taps each for SPLIT = 4; the
// quintic slices by Gauss point (t % SPLIT == sub). Each tap keeps the
// weight it has without the split.
template <int SHAPE, class R, int SPLIT = 1, class Interior, class Fast, class Visit>
__device__ __forceinline__ void segmentTaps(int nx, int ny, const int org[3], const R x[3], const R v[3], R first,
                                            R last, Interior interior, Fast fast, Visit visit, int sub = 0) {
  static_assert(SPLIT == 1 || SPLIT == 2 || SPLIT == 4, "tap slices: 1, 2 or 4");
  auto mine = [sub](int slice) { return SPLIT == 1 || slice % SPLIT == sub; };
  const R width = last - first;
  if (SHAPE == kShapeBinomial3X) {
    const R mid = (first + last) * R(0.5);
    const R px = x[0] + mid * v[0], py = x[1] + mid * v[1], pz = x[2] + mid * v[2];
    long long li = static_cast<long long>(floor(px)), lj = static_cast<long long>(floor(py));
    const long long lk = static_cast<long long>(floor(pz));
    long long i = li + org[0], j = lj + org[1];
    const long long k = lk + org[2];
    if (i == nx) { --i; --li; }
    if (j == ny) { --j; --lj; }
    R wx[4], wz[4];
    nodeWeights(px - static_cast<R>(li), wx);
    nodeWeights(pz - static_cast<R>(lk), wz);
    const R wy[2] = {R(1.) - (py - static_cast<R>(lj)), py - static_cast<R>(lj)};
    const R edge[3] = {R(0.25), R(0.5), R(0.25)};
    const R slope[4] = {R(-0.25), R(-0.25), R(0.25), R(0.25)};
    const R sy[2] = {R(-1.), R(1.)};
    const R curvature = width * width / R(12.);
    const R xy = curvature * v[0] * v[1];
    const R xz = curvature * v[0] * v[2];
    const R yz = curvature * v[1] * v[2];
    // binomialXZSegment visit order and weights, longitudinal_shape.h:81-93;
    // offsets are relative to (i-1, j, k-1)
    if (interior(i - 1, j, k - 1, 4, 2, 4)) {
      for (int c = 0; c < 4; ++c) {
        if (!mine(c)) continue;
        for (int b = 0; b < 2; ++b) {
          const R weight = width * (wy[b] * wz[c] + yz * sy[b] * slope[c]);
          for (int aa = 0; aa < 3; ++aa) fast(0, aa, b, c, weight * edge[aa]);
        }
        for (int aa = 0; aa < 4; ++aa) fast(1, aa, 0, c, width * (wx[aa] * wz[c] + xz * slope[aa] * slope[c]));
      }
      for (int b = 0; b < 2; ++b)
        for (int aa = 0; aa < 4; ++aa) {
          if (!mine(4 * b + aa)) continue;
          const R weight = width * (wx[aa] * wy[b] + xy * slope[aa] * sy[b]);
          for (int c = 0; c < 3; ++c) fast(2, aa, b, c, weight * edge[c]);
        }
      return;
    }
    for (int c = 0; c < 4; ++c) {
      if (!mine(c)) continue;
      for (int b = 0; b < 2; ++b) {
        const R weight = width * (wy[b] * wz[c] + yz * sy[b] * slope[c]);
        for (int aa = 0; aa < 3; ++aa) visit(0, i - 1 + aa, j + b, k - 1 + c, weight * edge[aa]);
      }
      for (int aa = 0; aa < 4; ++aa) visit(1, i - 1 + aa, j, k - 1 + c, width * (wx[aa] * wz[c] + xz * slope[aa] * slope[c]));
    }
    for (int b = 0; b < 2; ++b)
      for (int aa = 0; aa < 4; ++aa) {
        if (!mine(4 * b + aa)) continue;
        const R weight = width * (wx[aa] * wy[b] + xy * slope[aa] * sy[b]);
        for (int c = 0; c < 3; ++c) visit(2, i - 1 + aa, j + b, k - 1 + c, weight * edge[c]);
      }
  } else {
    for (int t = 0; t < 4; ++t) {
      if (!mine(t)) continue;
      const R node = static_cast<R>(gaussNode(t)), gw = static_cast<R>(gaussWeight(t));
      const R u = R(.5) * (last + first + width * node), factor = R(.5) * width * gw;
      const R xx = x[0] + u * v[0], yy = x[1] + u * v[1], zz = x[2] + u * v[2];
      long long li = static_cast<long long>(floor(xx)), lj = static_cast<long long>(floor(yy));
      long long i = li + org[0], j = lj + org[1];
      if (i == nx) { --i; --li; }
      if (j == ny) { --j; --lj; }
      const R fx = xx - static_cast<R>(li), fy = yy - static_cast<R>(lj);
      QBasisT<R> nodes, edge;
      quinticBasis(zz, false, nodes);
      quinticBasis(zz, true, edge);
      // offsets are relative to (i, j, first); edge.first == nodes.first
      const long long kn = static_cast<long long>(nodes.first) + org[2];
      const long long ke = static_cast<long long>(edge.first) + org[2];
      if (interior(i, j, kn, 2, 2, 12)) {
        for (int k = 0; k < 12; ++k) {
          const R w = factor * nodes.v[k];
          if (w == R(0.)) continue;
          for (int b = 0; b < 2; ++b) fast(0, 0, b, k, w * (b ? fy : R(1.) - fy));
          for (int a = 0; a < 2; ++a) fast(1, a, 0, k, w * (a ? fx : R(1.) - fx));
        }
        for (int k = 0; k < 11; ++k) {
          const R w = factor * edge.v[k];
          if (w == R(0.)) continue;
          for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a) fast(2, a, b, k, w * (a ? fx : R(1.) - fx) * (b ? fy : R(1.) - fy));
        }
        continue;
      }
      for (int k = 0; k < nodes.count; ++k) {
        const R w = factor * nodes.v[k];
        if (w == R(0.)) continue;
        for (int b = 0; b < 2; ++b) visit(0, i, j + b, kn + k, w * (b ? fy : R(1.) - fy));
        for (int a = 0; a < 2; ++a) visit(1, i + a, j, kn + k, w * (a ? fx : R(1.) - fx));
      }
      for (int k = 0; k < edge.count; ++k) {
        const R w = factor * edge.v[k];
        if (w == R(0.)) continue;
        for (int b = 0; b < 2; ++b)
          for (int a = 0; a < 2; ++a) visit(2, i + a, j + b, ke + k, w * (a ? fx : R(1.) - fx) * (b ? fy : R(1.) - fy));
      }
    }
  }
}

// Fixed-point tile partials (deposit side): the partial box of an occupied
// tile sits at its first chunk's slot, 3 components of boxNodes<SHAPE>()
// 64-bit integers scaled by 2^shift (shift per tile).
//
// Owner-side sum of lattice node (i, j, k) of component d of the tile
// (tx, ty, tz): over the neighbour tiles whose extended box covers the node,
// in ascending (z, y, x) tile order, the partial converted with that tile's
// exponent; each entry read is cleared for the next deposit. nbFirst /
// nbShift hold, per neighbour offset (oz + RZ, oy + 1, ox + 1), the first
// chunk slot (-1: unoccupied) and exponent.
template <int SHAPE>
__device__ __forceinline__ double ownedNodeSum(long long* partial, const int* nbFirst, const int* nbShift, int tx,
                                               int ty, int tz, int d, int i, int j, int k) {
  constexpr int N = boxNodes<SHAPE>(), NX = Box<SHAPE>::NX, NY = Box<SHAPE>::NY, NZ = Box<SHAPE>::NZ;
  constexpr int RZ = Box<SHAPE>::RZ;
  // tile t' covers node i when t'*T - X0 <= i < t'*T - X0 + NX
  auto lowTile = [](int v, int t0, int n, int tile) {
    const int a = v + t0 - n + tile;  // floor((v + t0 - n) / tile) + 1 for v + t0 - n >= -tile
    return a >= 0 ? a / tile : -((-a + tile - 1) / tile);
  };
  const int xLo = max(tx - 1, lowTile(i, Box<SHAPE>::X0, NX, kTileX)), xHi = min(tx + 1, (i + Box<SHAPE>::X0) / kTileX);
  const int yLo = max(ty - 1, lowTile(j, Box<SHAPE>::Y0, NY, kTileY)), yHi = min(ty + 1, (j + Box<SHAPE>::Y0) / kTileY);
  const int zLo = max(tz - RZ, lowTile(k, Box<SHAPE>::Z0, NZ, kTileZ)), zHi = min(tz + RZ, (k + Box<SHAPE>::Z0) / kTileZ);
  double sum = 0.0;
  for (int z = zLo; z <= zHi; ++z)
    for (int y = yLo; y <= yHi; ++y)
      for (int x = xLo; x <= xHi; ++x) {
        const int nb = ((z - tz + RZ) * 3 + (y - ty + 1)) * 3 + (x - tx + 1);
        const int f = nbFirst[nb];
        if (f < 0) continue;
        const int li = i - (x * kTileX - Box<SHAPE>::X0), lj = j - (y * kTileY - Box<SHAPE>::Y0),
                  lk = k - (z * kTileZ - Box<SHAPE>::Z0);
        long long* at = partial + static_cast<size_t>(f) * 3 * N + d * N + (lk * NY + lj) * NX + li;
        sum += ldexp(static_cast<double>(*at), -nbShift[nb]);
        *at = 0;
      }
  return sum;
}

// The owner side's table of the covering tiles of tile (tx, ty, tz):
// threads 0..9(2RZ+1)-1 fill, per neighbour offset, the first-chunk slot
// (-1: unoccupied) and the exponent from the tile scan (ownedNodeSum's
// nbFirst / nbShift). The caller synchronizes before reading it.
template <int SHAPE>
__device__ __forceinline__ void coveringTiles(const unsigned long long* tileScan, const int* tileShift, int tx,
                                              int ty, int tz, int ntx, int nty, int ntz, int* nbFirst,
                                              int* nbShift) {
  constexpr int RZ = Box<SHAPE>::RZ, NB = 9 * (2 * RZ + 1);
  if (threadIdx.x >= NB) return;
  const int nb = static_cast<int>(threadIdx.x);
  const int x = tx + nb % 3 - 1, y = ty + (nb / 3) % 3 - 1, z = tz + nb / 9 - RZ;
  int first = -1, sh = 0;
  if (x >= 0 && x < ntx && y >= 0 && y < nty && z >= 0 && z < ntz) {
    const int s = (z * nty + y) * ntx + x;
    const unsigned long long sc = tileScan[s];
    if (chunkField(tileScan[s + 1]) > chunkField(sc)) {
      first = occupiedField(sc);
      sh = tileShift[s];
    }
  }
  nbFirst[nb] = first;
  nbShift[nb] = sh;
}

// The deposit side's write-out of one chunk's LDS box acc (3 components):
// the tile's partial sits at its occupied-tile slot and holds zeros on
// entry (the reduction clears what it reads). A tile of one chunk stores
// its box; chunks of a larger tile add their nonzero nodes with integer
// atomics, whose sum is independent of their order.
template <int SHAPE>
__device__ __forceinline__ void flushTileBox(const unsigned long long* acc, long long* partial,
                                             const unsigned long long* tileScan, int tile) {
  constexpr int N = boxNodes<SHAPE>();
  const unsigned long long tsc = tileScan[tile];
  const bool alone = chunkField(tileScan[tile + 1]) - chunkField(tsc) == 1;
  unsigned long long* out =
      reinterpret_cast<unsigned long long*>(partial + static_cast<size_t>(occupiedField(tsc)) * 3 * N);
  for (int t = threadIdx.x; t < 3 * N; t += blockDim.x) {
    const unsigned long long val = acc[t];
    if (alone) out[t] = val;
    else if (val) atomicAdd(out + t, val);
  }
}

// Copy the tile box of E_cand + E_old into LDS (es, 3 components of
// boxNodes), zero outside the component ranges (such nodes are never read),
// with the workgroup's 2 kChunk lanes; every load of a lane is issued before
// its first LDS store. The component loop is unrolled so that its extents
// and base pointers are kernel-argument constants (a runtime component
// index would turn each of them into a dependent memory load). GA is the
// form's gather arguments (esum, ne), R its real type.
template <int SHAPE, class GA, class R>
__device__ void stageTileE(const GA& G, const TileBox<SHAPE>& box, R* es) {
  constexpr int N = boxNodes<SHAPE>();
  constexpr int kPer = (N + 2 * kChunk - 1) / (2 * kChunk);
  R ev[kPer][3];
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int q0 = static_cast<int>(threadIdx.x) + r * 2 * kChunk;
    const int li = q0 % Box<SHAPE>::NX, q = q0 / Box<SHAPE>::NX, lj = q % Box<SHAPE>::NY, lk = q / Box<SHAPE>::NY;
    const int i = box.x0 + li, j = box.y0 + lj, k = box.z0 + lk;
    const bool inBox = q0 < N && i >= 0 && j >= 0 && k >= 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      ev[r][d] = R(0);
      if (inBox && i < G.ne[d][0] && j < G.ne[d][1] && k < G.ne[d][2])
        ev[r][d] = G.esum[d][(static_cast<size_t>(k) * G.ne[d][1] + j) * G.ne[d][0] + i];
    }
  }
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int q0 = static_cast<int>(threadIdx.x) + r * 2 * kChunk;
    if (q0 < N)
#pragma unroll
      for (int d = 0; d < 3; ++d) es[d * N + q0] = ev[r][d];
  }
}

// A chunk's residual statistics in a fixed tree: the E lanes (threads
// 0..kChunk-1) enter (force, defect, scale, calls) of their particle; the
// chunk's max force and defect / scale / call sums go to out[0..3]. sh
// holds 4 kChunk doubles; every thread of the workgroup calls it.
__device__ __forceinline__ void chunkStatsTree(double* sh, bool enter, int lane, double force, double defect,
                                               double scale, double calls, double* out) {
  double* sf = sh;
  double* sd = sh + kChunk;
  double* ss = sh + 2 * kChunk;
  double* sc = sh + 3 * kChunk;
  if (enter) {
    sf[lane] = force;
    sd[lane] = defect;
    ss[lane] = scale;
    sc[lane] = calls;
  }
  __syncthreads();
  for (int w = kChunk / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) {
      sf[threadIdx.x] = fmax(sf[threadIdx.x], sf[threadIdx.x + w]);
      sd[threadIdx.x] += sd[threadIdx.x + w];
      ss[threadIdx.x] += ss[threadIdx.x + w];
      sc[threadIdx.x] += sc[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sf[0];
    out[1] = sd[0];
    out[2] = ss[0];
    out[3] = sc[0];
  }
}

}  // namespace aprl_gpu
