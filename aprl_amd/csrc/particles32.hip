// gfx950 particle kernels of the conservative-source step in the FP32-split
// form: fields, currents, gathered fields and momenta in FP32; positions as
// an absolute int32 lattice cell plus an FP32 in-cell offset.
//
// The step is the FP64 form's (tile sort, orbit gathers from LDS-staged
// tile boxes, Lapenta-Markidis push, fixed-point tile deposition, owner
// reduction) with these differences:
//  * shape weights, chord cuts and gathers run in FP32 on coordinates taken
//    relative to the particle's start cell, so the in-cell offset keeps its
//    24 bits whatever the absolute position;
//  * the push converts its inputs to FP64 and reuses the FP64 orbit push
//    (prescribed magnets, Newton root, boundary checks) on absolute box
//    coordinates, then stores the end point as cell + FP32 offset again;
//  * the Picard loop runs a fixed number of trials (at least the two the
//    reference requires with active particles, simulation_source_coupling.cpp:
//    50): its convergence tolerances (force < 1e-13) lie below FP32
//    resolution, and a fixed count keeps the host out of the step.

#include <algorithm>
#include <cmath>
#include <limits>

#include "diagnostics.hpp"
#include "particle_device.hpp"
#include "shapes.hpp"

namespace aprl_gpu {

// ---------------------------------------------------------------------------
// Kernel argument blocks.

struct GArgs32 {
  float* esum[3];         // E_cand + E_old of the trial
  const float* estar[3];  // zero-current preview E*
  const float* eold[3];   // E^n
  const float* b[3];      // B^(n+1/2)
  int ne[3][3], nb[3][3];
  int nx, ny, nz;
  int oc[3];       // absolute lattice cell of the grid origin
  float ch[3];     // c dt / h per axis (ballistic guess, cells per unit u/gamma)
  float negS;      // (-1*dt)/eps0
  int smoothX;
};

struct PArgs32 {
  const int* cell[3];
  const float* off[3];
  const float* u[3];
  const float* q;
  const float* m;
  float* fin[6];
  float* fout[6];
  int* cell1[3];
  float* off1[3];
  float* chord[3];  // trial chord x1 - x0 in cells, rounded once from FP64
  float* u1[3];
  float* gridAbs;
  const int* state;  // entry states (kStateActive / kStateCarrier for live entries)
  int* evt;          // boundary event code of the trial
  float* cut[3];     // t1, t2, carrier-part deposit weight
  float* wend;       // carrier end weight
  const float* w;    // weight slot (a carrier's current weight)
  const float* beta; // held speed of held entries
  const unsigned* tileOf;         // sorted tile key of each entry
  unsigned long long* tileBound;  // per tile: largest particle current bound
  float kq[3];  // h_a / (dx dy dz dt): current density per unit charge and cell of motion
  double h[3];
  int n;
};

// ---------------------------------------------------------------------------
// Chord of one particle relative to its start cell.

struct Chord32 {
  int org[3];        // grid index of the start cell
  float x[3], v[3];  // start offset and displacement, in cells
};

__device__ __forceinline__ Chord32 chordOf(const GArgs32& G, int c0x, int c0y, int c0z, const float off0[3],
                                           const float v[3]) {
  Chord32 r;
  r.org[0] = c0x - G.oc[0];
  r.org[1] = c0y - G.oc[1];
  r.org[2] = c0z - G.oc[2];
  for (int a = 0; a < 3; ++a) {
    r.x[a] = off0[a];
    r.v[a] = v[a];
  }
  return r;
}

// clipOpenOrbit (source_coupling_math.h:235-250) in start-cell coordinates:
// the grid spans [-org, n - org) per axis. On success the chord is replaced
// by its clipped part and the kept fraction is returned in kept.
__device__ bool clipChord(const GArgs32& G, Chord32& c, float& kept, unsigned& err) {
  float first = 0.f, last = 1.f;
  const int size[3] = {G.nx, G.ny, G.nz};
  bool valid = true;
  for (int a = 0; a < 3; ++a) {
    const float x = c.x[a], d = c.v[a];
    if (!isfinite(x) || !isfinite(d)) { err |= kErrNonfinite; return false; }
    const float lo = static_cast<float>(-c.org[a]) - x, hi = static_cast<float>(size[a] - c.org[a]) - x;
    if (d == 0.f) {
      if (lo > 0.f || hi < 0.f) valid = false;
      continue;
    }
    const float tlo = lo / d, thi = hi / d;
    if (d > 0.f) { first = fmaxf(first, tlo); last = fminf(last, thi); }
    else { last = fminf(last, tlo); first = fmaxf(first, thi); }
  }
  if (!valid || !(last >= first)) return false;
  for (int a = 0; a < 3; ++a) {
    const float d = c.v[a];
    if (first != 0.f) c.x[a] += first * d;
    c.v[a] = (last - first) * d;
  }
  kept = last - first;
  return true;
}

// ---------------------------------------------------------------------------
// Gathers from the staged tile box.

// Impl::electric (source_coupling.cpp:31-39): orbit average of
// 1/2 (E_cand + E_old) over the clipped chord with the deposit's weights.
template <int SHAPE>
__device__ void gatherE32(const GArgs32& G, const float* es, const TileBox<SHAPE>& box, Chord32 c, float res[3],
                          unsigned& err) {
  constexpr int N = boxNodes<SHAPE>(), NX = Box<SHAPE>::NX, NY = Box<SHAPE>::NY;
  res[0] = res[1] = res[2] = 0.f;
  float kept = 1.f;
  if (!clipChord(G, c, kept, err)) { err |= kErrOrbitOutside; return; }
  float cuts[kMaxCuts];
  const int count = orbitCuts(c.x, c.v, cuts, err);
  int base = 0;
  auto interior = [&](long long i, long long j, long long k, int wx, int wy, int wz) {
    if (i < 0 || j < 0 || k < 0 || i + wx > G.nx + 1 || j + wy > G.ny + 1 || k + wz > G.nz + 1) return false;
    if (SHAPE == kShapeBinomial3X) {
      if (i + 3 > G.nx || j + 1 > G.ny || k + 3 > G.nz) return false;
    } else {
      if (i + 1 > G.nx || j + 1 > G.ny || k + 11 > G.nz) return false;
    }
    if (!box.holds(i, j, k, wx, wy, wz)) return false;
    base = box.base(i, j, k);
    return true;
  };
  auto fast = [&](int d, int oi, int oj, int ok, float w) {
    res[d] += w * es[d * N + base + (ok * NY + oj) * NX + oi];
  };
  auto visit = [&](int d, long long ii, long long jj, long long kk, float w) {
    if (ii < 0 || jj < 0 || ii >= G.ne[d][0] || jj >= G.ne[d][1]) return;
    if (kk < 0 || kk >= G.ne[d][2]) return;
    const int l = box.local(static_cast<int>(ii), static_cast<int>(jj), static_cast<int>(kk), err);
    res[d] += w * (l < 0 ? 0.f : es[d * N + l]);
  };
  for (int s = 1; s < count; ++s) {
    if (cuts[s] - cuts[s - 1] <= 0.f) continue;
    segmentTaps<SHAPE, float>(G.nx, G.ny, c.org, c.x, c.v, cuts[s - 1], cuts[s], interior, fast, visit);
  }
  const float f = 0.5f * kept;
  res[0] *= f;
  res[1] *= f;
  res[2] *= f;
}

__device__ __forceinline__ int clampIndex32(int v, int n) { return v < 0 ? 0 : (v > n - 1 ? n - 1 : v); }

// Impl::magnetic for binomial3-x (source_coupling_math.h:142-167) at the
// clipped chord midpoint, clamped indices, read straight from B in HBM
// (the particles of a chunk share most of their 4x2x4 stencils in the
// caches), so that the gather stages only E.
__device__ void gatherBBinomial32(const GArgs32& G, Chord32 c, float res[3], unsigned& err) {
  res[0] = res[1] = res[2] = 0.f;
  float kept;
  if (!clipChord(G, c, kept, err)) { err |= kErrMagneticOutside; return; }
  float mid[3];
  for (int d = 0; d < 3; ++d) mid[d] = c.x[d] + 0.5f * c.v[d];
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    int cell[3];
    float frac[3];
    for (int d = 0; d < 3; ++d) {
      const float pos = mid[d] - (comp == d ? 0.f : 0.5f);
      const float fl = floorf(pos);
      cell[d] = static_cast<int>(fl) + c.org[d];
      frac[d] = pos - fl;
    }
    float wx[4], wz[4];
    nodeWeights(frac[0], wx);
    nodeWeights(frac[2], wz);
    const float wy[2] = {1.f - frac[1], frac[1]};
    const int n0 = G.nb[comp][0], n1 = G.nb[comp][1], n2 = G.nb[comp][2];
    const float* bc = G.b[comp];
    if (cell[0] >= 1 && cell[0] + 2 <= n0 - 1 && cell[1] >= 0 && cell[1] + 1 <= n1 - 1 && cell[2] >= 1 &&
        cell[2] + 2 <= n2 - 1) {
      const float* base = bc + (static_cast<size_t>(cell[2] - 1) * n1 + cell[1]) * n0 + (cell[0] - 1);
      for (int k = 0; k < 4; ++k)
        for (int j = 0; j < 2; ++j) {
          const float* r = base + (static_cast<size_t>(k) * n1 + j) * n0;
          float row = 0.f;
          for (int i = 0; i < 4; ++i) row += wx[i] * r[i];
          res[comp] += wz[k] * wy[j] * row;
        }
      continue;
    }
    for (int k = 0; k < 4; ++k)
      for (int j = 0; j < 2; ++j) {
        const int kk = clampIndex32(cell[2] - 1 + k, n2), jj = clampIndex32(cell[1] + j, n1);
        const float* r = bc + (static_cast<size_t>(kk) * n1 + jj) * n0;
        float row = 0.f;
        for (int i = 0; i < 4; ++i) row += wx[i] * r[clampIndex32(cell[0] - 1 + i, n0)];
        res[comp] += wz[k] * wy[j] * row;
      }
  }
}

// Impl::magnetic for the quintic shape (source_coupling_math.h:122-138):
// trilinear sample at the clipped chord midpoint, clamped indices, from HBM.
__device__ void gatherBQuintic32(const GArgs32& G, Chord32 c, float res[3], unsigned& err) {
  res[0] = res[1] = res[2] = 0.f;
  float kept;
  if (!clipChord(G, c, kept, err)) { err |= kErrMagneticNoDomain; return; }
  int cell[2][3];
  float fraction[2][3];
  for (int half = 0; half < 2; ++half)
    for (int d = 0; d < 3; ++d) {
      const float x = c.x[d] + 0.5f * c.v[d] - 0.5f * half;
      const float fl = floorf(x);
      cell[half][d] = static_cast<int>(fl) + c.org[d];
      fraction[half][d] = x - fl;
    }
#pragma unroll
  for (int axis = 0; axis < 3; ++axis) {
    const int hx = axis == 0 ? 0 : 1, hy = axis == 1 ? 0 : 1, hz = axis == 2 ? 0 : 1;
    const int* n = G.nb[axis];
    const float* src = G.b[axis];
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) {
          const float xw = i ? fraction[hx][0] : 1.f - fraction[hx][0];
          const float yw = j ? fraction[hy][1] : 1.f - fraction[hy][1];
          const float zw = k ? fraction[hz][2] : 1.f - fraction[hz][2];
          const int ii = clampIndex32(cell[hx][0] + i, n[0]), jj = clampIndex32(cell[hy][1] + j, n[1]),
                    kk = clampIndex32(cell[hz][2] + k, n[2]);
          res[axis] += xw * yw * zw * src[(static_cast<size_t>(kk) * n[1] + jj) * n[0] + ii];
        }
  }
}

// ---------------------------------------------------------------------------
// Sort.

__global__ void kTileKeys32(const int* cx, const int* cy, const int* cz, const float* oz, int* state, int n,
                            GArgs32 G, int ntx, int nty, unsigned* key, unsigned* perm, TileRangeArgs R, SortArgs K,
                            double hz) {
  __shared__ unsigned blockActive;
  if (threadIdx.x == 0) blockActive = 0u;
  __syncthreads();
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  clearTileRanges(R, p);
  if (p == 0) *K.removed = 0u;
  bool active = false;
  if (p < n) {
    perm[p] = p;
    const double z = (static_cast<double>(cz[p]) + static_cast<double>(oz[p])) * hz;
    if (!liveEntry(K, state, p, z)) {
      key[p] = static_cast<unsigned>(R.ntiles);
    } else {
      const int x = clampIndex32(cx[p] - G.oc[0], G.nx);
      const int y = clampIndex32(cy[p] - G.oc[1], G.ny);
      const int zc = clampIndex32(cz[p] - G.oc[2], G.nz);
      key[p] = (static_cast<unsigned>(zc / kTileZ) * nty + y / kTileY) * ntx + x / kTileX;
      active = state[p] == kStateActive;
    }
  }
  // live active entries: wave ballots summed in LDS, one global atomic per
  // block
  const unsigned long long act = __ballot(active);
  if (act && static_cast<int>(__lane_id()) == __ffsll(static_cast<long long>(act)) - 1)
    atomicAdd(&blockActive, static_cast<unsigned>(__popcll(act)));
  __syncthreads();
  if (threadIdx.x == 0 && blockActive) atomicAdd(K.words + 4, blockActive);
}

struct PermuteArgs32 {
  const int* cellSrc[3];
  int* cellDst[3];
  const float* src[9];
  float* dst[9];
  const uint64_t* idSrc;
  uint64_t* idDst;
  const int* stateSrc;
  int* stateDst;
  const float* betaSrc;
  float* betaDst;
};

// State scatter into sorted order, and the tile ranges of the sorted keys.
// blockIdx.y splits the gathered arrays over four times as many waves
// (y = 0: cells and offsets, tile ranges and bounds; 1: momenta; 2: q, m,
// w; 3: ids, states, held speeds).
__global__ void kPermute32(PermuteArgs32 A, const unsigned* perm, int n, TileBoundsArgs T) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int part = static_cast<int>(blockIdx.y);
  unsigned tile = ~0u;
  double bound = 0.0;
  if (p < n) {
    const unsigned from = perm[p];
    if (part < 3) {
#pragma unroll
      for (int a = 0; a < 3; ++a) A.dst[3 * part + a][p] = A.src[3 * part + a][from];
    } else {
      A.idDst[p] = A.idSrc[from];
      A.stateDst[p] = A.stateSrc[from];
      A.betaDst[p] = A.betaSrc[from];
    }
    if (part == 0) {
#pragma unroll
      for (int a = 0; a < 3; ++a) A.cellDst[a][p] = A.cellSrc[a][from];
      tileBoundsAt(T, p);
      const unsigned k = T.key[p];
      if (k < static_cast<unsigned>(T.ntiles)) {
        tile = k;
        bound = entryCurrentBound(static_cast<double>(A.src[6][from]), A.stateSrc[from], T.perCharge);
      }
    }
  }
  if (part == 0) raiseTileBound(T.bound, tile, bound);
}

// ---------------------------------------------------------------------------
// Gathers, push, deposit, reduction.

__device__ __forceinline__ void chordAt(const PArgs32& P, int p, bool guess, const GArgs32& G, int c0[3],
                                        float off0[3], float v[3]) {
  for (int d = 0; d < 3; ++d) {
    c0[d] = P.cell[d][p];
    off0[d] = P.off[d][p];
  }
  if (guess) {
    // ballistic end point x0 + c dt u0 / gamma0 (source_coupling.cpp:133-147)
    const float ux = P.u[0][p], uy = P.u[1][p], uz = P.u[2][p];
    const float ig = rsqrtf(1.f + ux * ux + uy * uy + uz * uz);
    v[0] = G.ch[0] * ux * ig;
    v[1] = G.ch[1] * uy * ig;
    v[2] = G.ch[2] * uz * ig;
  } else {
    // the push's displacement rounded once: keeps its relative precision
    // where x1 - x0 is far below the offsets' resolution (transverse motion)
    for (int d = 0; d < 3; ++d) v[d] = P.chord[d][p];
  }
}

// Chunked orbit gathers (one workgroup of 2*kChunk lanes per chunk: E
// lanes, then B lanes). MODE 0 gathers at (x0, ballistic guess) against the
// zero-current preview into fin; MODE 1 is the residual gather at (x0, x1)
// into fout, with the chunk's force max and defect / scale / call sums when
// chunkOut is given.
template <int MODE, int SHAPE>
__global__ void __launch_bounds__(2 * kChunk) kGatherChunks32(PArgs32 P, GArgs32 G, StepArgs S, const int4* chunks,
                                                              int ntx, int nty, double dt, double* chunkOut,
                                                              unsigned* errOut) {
  __shared__ float es[3 * boxNodes<SHAPE>()];
  __shared__ double red[MODE == 1 ? 4 * kChunk : 1];
  const int cb = xcdContiguous(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
  const int4 ch = chunks[cb];
  const TileBox<SHAPE> box = tileBox<SHAPE>(ch.x, ntx, nty);
  stageTileE<SHAPE>(G, box, es);
  __syncthreads();
  const bool bWaves = threadIdx.x >= kChunk;
  const int lane = static_cast<int>(threadIdx.x) - (bWaves ? kChunk : 0);
  const int p = ch.y + lane;
  // carriers move without fields and take no part in the residual; a held
  // entry takes part from its release point with the rest of the step
  const int st = p < ch.z ? P.state[p] : kStateRemoved;
  bool live = st == kStateActive, released = false;
  double dtp = dt;
  float dk[3] = {0.f, 0.f, 0.f};  // held leg to the release point, in cells
  if (st == kStateHeld) {
    double x0[3], uh[3];
    for (int d = 0; d < 3; ++d) {
      x0[d] = (static_cast<double>(P.cell[d][p]) + static_cast<double>(P.off[d][p])) * P.h[d];
      uh[d] = P.u[d][p];
    }
    const HeldMove h = heldMove(S, x0, uh, static_cast<double>(P.beta[p]));
    if (h.release && h.remaining > 0.0) {
      live = released = true;
      dtp = h.remaining;
      for (int d = 0; d < 3; ++d) dk[d] = static_cast<float>((h.x1[d] - x0[d]) / P.h[d]);
    }
  }
  unsigned err = 0;
  float g[3] = {0.f, 0.f, 0.f};
  int c0[3] = {0, 0, 0};
  float off0[3] = {0.f, 0.f, 0.f}, v[3] = {0.f, 0.f, 0.f};
  if (live) {
    chordAt(P, p, MODE == 0, G, c0, off0, v);
    if (released) {
      for (int d = 0; d < 3; ++d) off0[d] += dk[d];
      if (MODE == 0) {
        const float r = static_cast<float>(dtp / dt);
        for (int d = 0; d < 3; ++d) v[d] *= r;
      } else {
        for (int d = 0; d < 3; ++d) v[d] -= dk[d];
      }
    }
    const Chord32 c = chordOf(G, c0[0], c0[1], c0[2], off0, v);
    if (bWaves) {
      if constexpr (SHAPE == kShapeBinomial3X) gatherBBinomial32(G, c, g, err);
      else gatherBQuintic32(G, c, g, err);
    } else {
      gatherE32<SHAPE>(G, es, box, c, g, err);
    }
    float* const* dst = MODE == 0 ? P.fin : P.fout;
    const int off = bWaves ? 3 : 0;
    for (int d = 0; d < 3; ++d) dst[off + d][p] = g[d];
  }
  if (MODE == 1 && chunkOut) {
    // localResidual (source_coupling.cpp:105-159) per chunk, fixed tree order
    __syncthreads();
    float* xb = es;  // B of the B lanes, for the E lanes
    if (bWaves)
      for (int d = 0; d < 3; ++d) xb[d * kChunk + lane] = g[d];
    __syncthreads();
    double force = 0., defect = 0., scale = 0., calls = 0.;
    if (!bWaves && live) {
      double de[3], db[3];
      for (int d = 0; d < 3; ++d) {
        de[d] = static_cast<double>(g[d]) - P.fin[d][p];
        db[d] = static_cast<double>(xb[d * kChunk + lane]) - P.fin[3 + d][p];
      }
      const double q = P.q[p], m = P.m[p];
      force = fabs(q / m) * dtp / kC * (sqrt(de[0] * de[0] + de[1] * de[1] + de[2] * de[2]) +
                                        kC * sqrt(db[0] * db[0] + db[1] * db[1] + db[2] * db[2]));
      double w1 = 0., w2 = 0.;
      for (int d = 0; d < 3; ++d) {
        const double dx = static_cast<double>(v[d]) * P.h[d];
        w1 += q * de[d] * dx;
        w2 += q * g[d] * dx;
      }
      defect = fabs(w1);
      scale = static_cast<double>(P.gridAbs[p]) + fabs(w2);
      calls = 1.;
    }
    chunkStatsTree(red, !bWaves, lane, force, defect, scale, calls, chunkOut + 4 * cb);
  }
  if (err) atomicOr(errOut, err);
}

// One entry of kPushSplit; returns the entry's current bound: the largest
// |scale_a| of its FP32 chord as the deposit forms it (q kq_a v_a), plus,
// for a held entry whose release kink splits the path, the held leg's
// largest scale.
__device__ __forceinline__ void pushSplitEntry(const PArgs32& P, const StepArgs& S, double o0, double o1, double o2,
                                               int p, unsigned* errOut) {
  unsigned err = 0;
  const double o[3] = {o0, o1, o2};
  int c0[3];
  double x0[3], u0[3];
  for (int a = 0; a < 3; ++a) {
    c0[a] = P.cell[a][p];
    x0[a] = (static_cast<double>(c0[a]) + static_cast<double>(P.off[a][p])) * P.h[a];
    u0[a] = P.u[a][p];
  }
  double x[3], u[3], absGrid, t1, t2, wDeposit, wEnd, usedE[3], usedB[3];
  for (int a = 0; a < 3; ++a) {
    usedE[a] = P.fin[a][p];
    usedB[a] = P.fin[3 + a][p];
  }
  const int code = entryStep(S, o, P.state[p], x0, u0, P.q[p], P.m[p], static_cast<double>(P.w[p]),
                             static_cast<double>(P.beta[p]), usedE, usedB, x, u, absGrid, t1, t2, wDeposit, wEnd,
                             err);
  for (int a = 0; a < 3; ++a) {
    const double rel = static_cast<double>(P.off[a][p]) + (x[a] - x0[a]) / P.h[a];
    double fl = floor(rel);
    float off = static_cast<float>(rel - fl);
    if (off >= 1.f) {  // rel - fl rounded up to 1
      off = 0.f;
      fl += 1.;
    }
    int cell = c0[a];
    if (isfinite(fl) && fabs(fl) < 1e9) cell += static_cast<int>(fl);
    else err |= kErrNonfinite;
    P.cell1[a][p] = cell;
    P.off1[a][p] = off;
    P.chord[a][p] = static_cast<float>((x[a] - x0[a]) / P.h[a]);
    P.u1[a][p] = static_cast<float>(u[a]);
  }
  P.gridAbs[p] = static_cast<float>(absGrid);
  P.evt[p] = code;
  if (code) {
    P.cut[0][p] = static_cast<float>(t1);
    P.cut[1][p] = static_cast<float>(t2);
    P.cut[2][p] = static_cast<float>(wDeposit);
    P.wend[p] = static_cast<float>(wEnd);
    err |= kEvtAny;
  }
  if (err) atomicOr(errOut, err);
}

// One trial's push: FP32 state and fields widened to FP64, the FP64 orbit
// push on absolute box coordinates (origin o = the grid origin cell times
// h) or a carrier's straight move (position += dt u c / gamma), the
// segment's boundary events, and the end point split again into cell +
// FP32 offset; then the entries' tile current bounds.
__global__ void __launch_bounds__(128) kPushSplit(PArgs32 P, StepArgs S, double o0, double o1, double o2,
                                                  unsigned* errOut) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < P.n) pushSplitEntry(P, S, o0, o1, o2, p, errOut);
}

// Deterministic tile deposition (Current::deposit, longitudinal_shape.h:
// 175-187) of the FP32 chords into 64-bit fixed-point LDS tile boxes; the
// scheme is the FP64 form's: one exponent per tile from the tile's current
// bound, integer sums independent of order.
template <int SHAPE>
__global__ void __launch_bounds__(kChunk * depositLanes<SHAPE>()) kDepositSplit(PArgs32 P, GArgs32 G, StepArgs S, const int4* chunks,
                                                        const int* tileBegin,
                                                        const int* tileEnd, const unsigned long long* tileScan,
                                                        int ntx, int nty, long long* partial, int* tileShift,
                                                        unsigned* errOut) {
  constexpr int N = boxNodes<SHAPE>(), NX = Box<SHAPE>::NX, NY = Box<SHAPE>::NY;
  __shared__ unsigned long long acc[3 * N];
  const int4 ch = chunks[blockIdx.x];
  const int tile = ch.x;
  const TileBox<SHAPE> box = tileBox<SHAPE>(tile, ntx, nty);
  for (int t = threadIdx.x; t < 3 * N; t += blockDim.x) acc[t] = 0ull;
  // scale_a = q * dx_a / (dx dy dz dt), eb_deposition.cpp:41-51
  auto scaleOf = [&](int p, const float v[3], float s[3]) {
    const float q = P.q[p];
    for (int d = 0; d < 3; ++d) s[d] = q * P.kq[d] * v[d];
  };
  const unsigned long long tsc = tileScan[tile];
  const int shift = tileFixedShift(Box<SHAPE>::kWeightBound, P.tileBound[tile], tileEnd[tile] - tileBegin[tile]);
  if (threadIdx.x == 0 && static_cast<int>(blockIdx.x) == chunkField(tsc)) tileShift[tile] = shift;
  __syncthreads();
  unsigned err = 0;
  // lanes [s * kChunk, (s + 1) * kChunk) take tap slice s of particles ch.y + lane
  const int sub = static_cast<int>(threadIdx.x) / kChunk;
  const int p = ch.y + static_cast<int>(threadIdx.x) % kChunk;
  if (p < ch.z) {
    int c0[3];
    float off0[3], v[3], scale[3];
    chordAt(P, p, false, G, c0, off0, v);
    // A held entry released inside the step deposits its held leg to the
    // release point first, then the rest of the chord.
    float dk[3] = {0.f, 0.f, 0.f};
    int legs = 1;
    if (P.state[p] == kStateHeld) {
      double x0[3], uh[3];
      for (int d = 0; d < 3; ++d) {
        x0[d] = (static_cast<double>(c0[d]) + static_cast<double>(off0[d])) * P.h[d];
        uh[d] = P.u[d][p];
      }
      const HeldMove h = heldMove(S, x0, uh, static_cast<double>(P.beta[p]));
      if (h.release && h.kink) {
        legs = 2;
        for (int d = 0; d < 3; ++d) dk[d] = static_cast<float>((h.x1[d] - x0[d]) / P.h[d]);
      }
    }
    Chord32 c = chordOf(G, c0[0], c0[1], c0[2], off0, v);
    float cuts[kMaxCuts];
    int base = 0;
    auto interior = [&](long long i, long long j, long long k, int wx, int wy, int wz) {
      if (i < 0 || j < 0 || k < 0 || i + wx > G.nx + 1 || j + wy > G.ny + 1 || k + wz > G.nz + 1) return false;
      if (SHAPE == kShapeBinomial3X) {
        if (i + 3 > G.nx || j + 1 > G.ny || k + 3 > G.nz) return false;
      } else {
        if (i + 1 > G.nx || j + 1 > G.ny || k + 11 > G.nz) return false;
      }
      if (!box.holds(i, j, k, wx, wy, wz)) return false;
      base = box.base(i, j, k);
      return true;
    };
    auto fixedOf = [&](float value) {
      return static_cast<unsigned long long>(llrint(ldexp(static_cast<double>(value), shift)));
    };
    float mult = 1.f;  // weight of the chord part being deposited
    auto fast = [&](int d, int oi, int oj, int ok, float w) {
      if (scale[d] == 0.f) return;
      atomicAdd(&acc[d * N + base + (ok * NY + oj) * NX + oi], fixedOf(scale[d] * w * mult));
    };
    auto visit = [&](int d, long long ii, long long jj, long long kk, float w) {
      if (scale[d] == 0.f) return;
      if (ii < 0 || ii >= G.ne[d][0]) {
        if (SHAPE != kShapeBinomial3X || !G.smoothX) err |= kErrCurrentX;
        return;
      }
      if (jj < 0 || jj >= G.ne[d][1]) { err |= kErrCurrentY; return; }
      if (kk < 0 || kk >= G.ne[d][2]) return;
      const int l = box.local(static_cast<int>(ii), static_cast<int>(jj), static_cast<int>(kk), err);
      if (l < 0) return;
      atomicAdd(&acc[d * N + l], fixedOf(scale[d] * w * mult));
    };
    // A boundary event splits the last leg: [0, t1] at unit weight (the
    // active part) and [t1, t2] at the carrier part's deposit weight.
    const int e = P.evt[p];
    const float t1 = e ? P.cut[0][p] : 1.f, t2 = e ? P.cut[1][p] : 1.f, wCarrier = e ? P.cut[2][p] : 0.f;
#pragma unroll 1
    for (int leg = 0; leg < legs; ++leg) {
      const bool last = leg + 1 == legs;
      if (legs == 2) {
        // leg 0: dk from the start; leg 1: the rest from the kink
        for (int d = 0; d < 3; ++d) {
          c.x[d] = off0[d] + (leg ? dk[d] : 0.f);
          c.v[d] = leg ? v[d] - dk[d] : dk[d];
        }
      }
      scaleOf(p, c.v, scale);
      const int count = orbitCuts(c.x, c.v, cuts, err);
      const float f1 = last ? t1 : 1.f, f2 = last ? t2 : 1.f, fw = last ? wCarrier : 0.f;
      for (int sgm = 1; sgm < count; ++sgm) {
        const float k0 = cuts[sgm - 1], k1 = cuts[sgm];
        if (k1 - k0 <= 0.f) continue;
        // without an event: one part [k0, k1] (t1 = t2 = 1)
#pragma unroll 1
        for (int part = 0; part < 2; ++part) {
          const float a = part ? fmaxf(k0, f1) : k0;
          const float b = part ? fminf(k1, f2) : fminf(k1, f1);
          if (!(b > a)) continue;
          mult = part ? fw : 1.f;
          segmentTaps<SHAPE, float, depositLanes<SHAPE>()>(G.nx, G.ny, c.org, c.x, c.v, a, b, interior, fast, visit, sub);
        }
      }
    }
  }
  __syncthreads();
  flushTileBox<SHAPE>(acc, partial, tileScan, tile);
  if (err) atomicOr(errOut, err);
}

// Node-owner reduction of the tile partials into FP32 J (sums in FP64, one
// rounding), clearing what it read; with WRITE_ESUM it also writes
// E_cand + E_old (applyCurrentSource(-1) on the preview,
// source_coupling.cpp:84-89) for the next residual gather.
template <int SHAPE, bool WRITE_ESUM>
__global__ void __launch_bounds__(256) kReduceCurrentSplit(float* jx, float* jy, float* jz, GArgs32 G,
                                                           const int* active, const unsigned long long* tileScan,
                                                           const int* tileShift, long long* partial, int ntx,
                                                           int nty, int ntz) {
  constexpr int RZ = Box<SHAPE>::RZ, NB = 9 * (2 * RZ + 1);
  const int tile = active[xcdContiguous(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x))];
  const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
  __shared__ int nbFirst[NB], nbShift[NB];
  coveringTiles<SHAPE>(tileScan, tileShift, tx, ty, tz, ntx, nty, ntz, nbFirst, nbShift);
  __syncthreads();
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int n0 = G.ne[d][0], n1 = G.ne[d][1], n2 = G.ne[d][2];
    const int i0 = tx * kTileX, i1 = tx == ntx - 1 ? n0 : min(n0, i0 + kTileX);
    const int j0 = ty * kTileY, j1 = ty == nty - 1 ? n1 : min(n1, j0 + kTileY);
    const int k0 = tz * kTileZ, k1 = tz == ntz - 1 ? n2 : min(n2, k0 + kTileZ);
    const int wx = i1 - i0, wy = j1 - j0, wz = k1 - k0;
    float* dst = d == 0 ? jx : (d == 1 ? jy : jz);
    for (int t = threadIdx.x; t < wx * wy * wz; t += blockDim.x) {
      const int i = i0 + t % wx, j = j0 + (t / wx) % wy, k = k0 + t / (wx * wy);
      const double sum = ownedNodeSum<SHAPE>(partial, nbFirst, nbShift, tx, ty, tz, d, i, j, k);
      const size_t idx = (static_cast<size_t>(k) * n1 + j) * n0 + i;
      const float jv = static_cast<float>(sum);
      dst[idx] = jv;
      if (WRITE_ESUM) {
        const bool applied = d == 0 ? (j > 0 && j < G.ny && k != 0 && k != G.nz)
                                    : (d == 1 ? (i > 0 && i < G.nx && k != 0 && k != G.nz)
                                              : (i > 0 && i < G.nx && j > 0 && j < G.ny));
        float e = G.estar[d][idx];
        if (applied && jv != 0.f) e += G.negS * jv;
        G.esum[d][idx] = e + G.eold[d][idx];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Host side.

void reserveSplitParticles(SplitStore& P, std::size_t capacity) {
  if (capacity <= P.capacity) return;
  for (int a = 0; a < 3; ++a) {
    P.cell[a].resize(capacity);
    P.cellT[a].resize(capacity);
    P.endCell[a].resize(capacity);
    P.chord[a].resize(capacity);
  }
  for (int a = 0; a < 9; ++a) {
    P.s[a].resize(capacity);
    P.t[a].resize(capacity);
  }
  P.id.resize(capacity);
  P.idT.resize(capacity);
  P.state.resize(capacity);
  P.stateT.resize(capacity);
  P.beta.resize(capacity);
  P.betaT.resize(capacity);
  P.evt.resize(capacity);
  for (int a = 0; a < 3; ++a) P.cut[a].resize(capacity);
  P.wend.resize(capacity);
  if (!P.removed.size()) P.removed.resize(1);
  for (int a = 0; a < 6; ++a) {
    P.fields[0][a].resize(capacity);
    P.fields[1][a].resize(capacity);
    P.end[a].resize(capacity);
  }
  P.gridAbs.resize(capacity);
  P.key.resize(capacity);
  P.keySorted.resize(capacity);
  P.perm.resize(capacity);
  P.permSorted.resize(capacity);
  P.capacity = capacity;
  P.sortTempBytes = 0;
  tileSort(nullptr, P.sortTempBytes, P.key.get(), P.keySorted.get(), P.perm.get(), P.permSorted.get(), static_cast<unsigned>(capacity), 32, nullptr);
  P.sortTempBytes = std::max<std::size_t>(P.sortTempBytes, 256);
  P.sortTemp.resize(P.sortTempBytes);
}

namespace {
// Split absolute coordinate x (metres) on the lattice of spacing h.
inline void splitCoordinate(double x, double h, int& cell, float& off) {
  const double r = x / h;
  if (!std::isfinite(r) || std::fabs(r) >= 2.0e9)
    throw std::out_of_range("Particle position is outside the FP32-split lattice range");
  double fl = std::floor(r);
  float f = static_cast<float>(r - fl);
  if (f >= 1.0f) {
    f = 0.0f;
    fl += 1.0;
  }
  cell = static_cast<int>(fl);
  off = f;
}
}  // namespace

void uploadSplitParticles(SplitStore& P, const FieldState32& F, std::size_t n, const double* x3, const double* u3,
                          const double* q, const double* m, const double* w, const uint64_t* id) {
  reserveSplitParticles(P, n);
  const double h[3] = {F.g.dx, F.g.dy, F.g.dz};
  std::vector<int> cells(n);
  std::vector<float> col(n);
  for (int a = 0; a < 3; ++a) {
    for (std::size_t i = 0; i < n; ++i) splitCoordinate(x3[3 * i + a], h[a], cells[i], col[i]);
    hipCheck(hipMemcpy(P.cell[a].get(), cells.data(), n * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
    hipCheck(hipMemcpy(P.s[a].get(), col.data(), n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
    for (std::size_t i = 0; i < n; ++i) col[i] = static_cast<float>(u3[3 * i + a]);
    hipCheck(hipMemcpy(P.s[3 + a].get(), col.data(), n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
  }
  const double* scalars[3] = {q, m, w};
  for (int a = 0; a < 3; ++a) {
    for (std::size_t i = 0; i < n; ++i) col[i] = static_cast<float>(scalars[a][i]);
    hipCheck(hipMemcpy(P.s[6 + a].get(), col.data(), n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
  }
  hipCheck(hipMemcpy(P.id.get(), id, n * sizeof(uint64_t), hipMemcpyHostToDevice), "hipMemcpy");
  P.count = n;
  ++P.generation;
}

// Grows the store keeping the persistent state of its first `count`
// entries (cells, offsets, momenta, charges, ids, entry states, held
// speeds); the per-step arrays are rewritten by every step.
void growSplitParticles(SplitStore& P, std::size_t capacity) {
  if (capacity <= P.capacity) return;
  DeviceArray<int> cell[3], state;
  DeviceArray<float> s[9], beta;
  DeviceArray<uint64_t> id;
  for (int a = 0; a < 3; ++a) cell[a].swap(P.cell[a]);
  for (int a = 0; a < 9; ++a) s[a].swap(P.s[a]);
  id.swap(P.id);
  state.swap(P.state);
  beta.swap(P.beta);
  reserveSplitParticles(P, capacity);
  const std::size_t n = P.count;
  if (!n) return;
  for (int a = 0; a < 3; ++a)
    hipCheck(hipMemcpy(P.cell[a].get(), cell[a].get(), n * sizeof(int), hipMemcpyDeviceToDevice), "hipMemcpy");
  for (int a = 0; a < 9; ++a)
    hipCheck(hipMemcpy(P.s[a].get(), s[a].get(), n * sizeof(float), hipMemcpyDeviceToDevice), "hipMemcpy");
  hipCheck(hipMemcpy(P.id.get(), id.get(), n * sizeof(uint64_t), hipMemcpyDeviceToDevice), "hipMemcpy");
  hipCheck(hipMemcpy(P.state.get(), state.get(), n * sizeof(int), hipMemcpyDeviceToDevice), "hipMemcpy");
  hipCheck(hipMemcpy(P.beta.get(), beta.get(), n * sizeof(float), hipMemcpyDeviceToDevice), "hipMemcpy");
}

void appendSplitParticles(SplitStore& P, const FieldState32& F, std::size_t k, const double* x3, const double* u3,
                          const double* q, const double* m, const double* w, const uint64_t* id, const int* state,
                          const double* beta) {
  if (!k) return;
  const std::size_t n = P.count;
  const double h[3] = {F.g.dx, F.g.dy, F.g.dz};
  std::vector<float> col(k);
  // convert before growing, so that a rejected position leaves the store as it was
  std::vector<int> allCells(3 * k);
  std::vector<float> offsets(3 * k);
  for (int a = 0; a < 3; ++a)
    for (std::size_t i = 0; i < k; ++i) splitCoordinate(x3[3 * i + a], h[a], allCells[a * k + i], offsets[a * k + i]);
  if (n + k > P.capacity) growSplitParticles(P, std::max(n + k, P.capacity + P.capacity / 4 + 1024));
  for (int a = 0; a < 3; ++a) {
    hipCheck(hipMemcpy(P.cell[a].get() + n, allCells.data() + a * k, k * sizeof(int), hipMemcpyHostToDevice),
             "hipMemcpy");
    hipCheck(hipMemcpy(P.s[a].get() + n, offsets.data() + a * k, k * sizeof(float), hipMemcpyHostToDevice),
             "hipMemcpy");
    for (std::size_t i = 0; i < k; ++i) col[i] = static_cast<float>(u3[3 * i + a]);
    hipCheck(hipMemcpy(P.s[3 + a].get() + n, col.data(), k * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
  }
  const double* scalars[4] = {q, m, w, beta};
  DeviceArray<float>* dst[4] = {&P.s[6], &P.s[7], &P.s[8], &P.beta};
  for (int a = 0; a < 4; ++a) {
    for (std::size_t i = 0; i < k; ++i) col[i] = static_cast<float>(scalars[a][i]);
    hipCheck(hipMemcpy(dst[a]->get() + n, col.data(), k * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
  }
  hipCheck(hipMemcpy(P.id.get() + n, id, k * sizeof(uint64_t), hipMemcpyHostToDevice), "hipMemcpy");
  hipCheck(hipMemcpy(P.state.get() + n, state, k * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
  P.count = n + k;
  ++P.generation;
}

void downloadSplitParticles(const SplitStore& P, const FieldState32& F, double* x3, double* u3, double* q, double* m,
                            double* w, uint64_t* id) {
  hipCheck(hipStreamSynchronize(F.stream), "hipStreamSynchronize");
  const std::size_t n = P.count;
  const double h[3] = {F.g.dx, F.g.dy, F.g.dz};
  std::vector<int> cells(n);
  std::vector<float> col(n);
  for (int a = 0; a < 3; ++a) {
    if (x3) {
      hipCheck(hipMemcpy(cells.data(), P.cell[a].get(), n * sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy");
      hipCheck(hipMemcpy(col.data(), P.s[a].get(), n * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
      for (std::size_t i = 0; i < n; ++i)
        x3[3 * i + a] = (static_cast<double>(cells[i]) + static_cast<double>(col[i])) * h[a];
    }
    if (u3) {
      hipCheck(hipMemcpy(col.data(), P.s[3 + a].get(), n * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
      for (std::size_t i = 0; i < n; ++i) u3[3 * i + a] = col[i];
    }
  }
  double* scalars[3] = {q, m, w};
  for (int a = 0; a < 3; ++a)
    if (scalars[a]) {
      hipCheck(hipMemcpy(col.data(), P.s[6 + a].get(), n * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
      for (std::size_t i = 0; i < n; ++i) scalars[a][i] = col[i];
    }
  if (id) hipCheck(hipMemcpy(id, P.id.get(), n * sizeof(uint64_t), hipMemcpyDeviceToHost), "hipMemcpy");
}

namespace {

GArgs32 gatherArgs32(CouplingState& C, const FieldState32& F, const int oc[3]) {
  GArgs32 G{};
  for (int a = 0; a < 3; ++a) {
    G.esum[a] = C.esum32[a].get();
    G.estar[a] = C.estar32[a].get();
    G.eold[a] = F.base(a);
    G.b[a] = F.base(3 + a);
    for (int d = 0; d < 3; ++d) {
      G.ne[a][d] = static_cast<int>(F.dims[a][d]);
      G.nb[a][d] = static_cast<int>(F.dims[3 + a][d]);
    }
    G.oc[a] = oc[a];
  }
  G.nx = static_cast<int>(F.g.nx);
  G.ny = static_cast<int>(F.g.ny);
  G.nz = static_cast<int>(F.g.nz);
  G.ch[0] = static_cast<float>(F.g.dt * kC / F.g.dx);
  G.ch[1] = static_cast<float>(F.g.dt * kC / F.g.dy);
  G.ch[2] = static_cast<float>(F.g.dt * kC / F.g.dz);
  G.negS = static_cast<float>(-1.0 * F.g.dt / kEps0);
  G.smoothX = C.smoothX ? 1 : 0;
  return G;
}

PArgs32 particleArgs32(SplitStore& P, const FieldState32& F) {
  PArgs32 A{};
  for (int a = 0; a < 3; ++a) {
    A.cell[a] = P.cell[a].get();
    A.off[a] = P.s[a].get();
    A.u[a] = P.s[3 + a].get();
    A.cell1[a] = P.endCell[a].get();
    A.off1[a] = P.end[a].get();
    A.chord[a] = P.chord[a].get();
    A.u1[a] = P.end[3 + a].get();
  }
  for (int a = 0; a < 6; ++a) {
    A.fin[a] = P.fields[0][a].get();
    A.fout[a] = P.fields[1][a].get();
  }
  A.q = P.s[6].get();
  A.m = P.s[7].get();
  A.w = P.s[8].get();
  A.beta = P.beta.get();
  A.gridAbs = P.gridAbs.get();
  A.state = P.state.get();
  A.evt = P.evt.get();
  for (int a = 0; a < 3; ++a) A.cut[a] = P.cut[a].get();
  A.wend = P.wend.get();
  const double h[3] = {F.g.dx, F.g.dy, F.g.dz};
  const double den = F.g.dx * F.g.dy * F.g.dz * F.g.dt;
  for (int a = 0; a < 3; ++a) {
    A.h[a] = h[a];
    A.kq[a] = static_cast<float>(h[a] / den);
  }
  A.n = static_cast<int>(P.count);
  return A;
}

int bitsFor32(unsigned v) {
  int b = 1;
  while (b < 32 && (v >> b) != 0u) ++b;
  return b;
}

void sortSplit(CouplingState& C, SplitStore& P, const GArgs32& G, const SortArgs& K, int ntx, int nty, int ntiles,
               hipStream_t st) {
  FieldState32& F = *C.fields32;
  const int n = static_cast<int>(P.count);
  ScopedKernel sk(F.profiler, st, "sort");
  const unsigned blocks = ceilDiv(n, 256);
  kTileKeys32<<<blocks, 256, 0, st>>>(P.cell[0].get(), P.cell[1].get(), P.cell[2].get(), P.s[2].get(),
                                      P.state.get(), n, G, ntx, nty, P.key.get(), P.perm.get(),
                                      TileRangeArgs{C.tileBegin.get(), C.tileEnd.get(), ntiles, C.tileBound.get()}, K,
                                      F.g.dz);
  launchCheck("tile keys");
  const int bits = bitsFor32(static_cast<unsigned>(ntiles));  // keys 0..ntiles (ntiles: behind the live entries)
  std::size_t bytes = 0;
  tileSort(nullptr, bytes, P.key.get(), P.keySorted.get(), P.perm.get(), P.permSorted.get(), static_cast<unsigned>(n), bits, st);
  if (bytes > P.sortTempBytes) {
    P.sortTemp.resize(bytes);
    P.sortTempBytes = bytes;
  }
  bytes = P.sortTempBytes;
  tileSort(P.sortTemp.get(), bytes, P.key.get(), P.keySorted.get(), P.perm.get(), P.permSorted.get(), static_cast<unsigned>(n), bits, st);
  PermuteArgs32 A{};
  for (int a = 0; a < 3; ++a) {
    A.cellSrc[a] = P.cell[a].get();
    A.cellDst[a] = P.cellT[a].get();
  }
  for (int a = 0; a < 9; ++a) {
    A.src[a] = P.s[a].get();
    A.dst[a] = P.t[a].get();
  }
  A.idSrc = P.id.get();
  A.idDst = P.idT.get();
  A.stateSrc = P.state.get();
  A.stateDst = P.stateT.get();
  A.betaSrc = P.beta.get();
  A.betaDst = P.betaT.get();
  // |q kq_a v_a| with |v_a| < c dt / h_a cells, with a margin for the FP32 rounding
  const double h[3] = {F.g.dx, F.g.dy, F.g.dz};
  double perCharge = 0.0;
  for (int a = 0; a < 3; ++a)
    perCharge = std::max(perCharge, static_cast<double>(static_cast<float>(h[a] / (F.g.dx * F.g.dy * F.g.dz * F.g.dt))) *
                                        (kC * F.g.dt / h[a]));
  perCharge *= 1.0 + 1e-6;
  const TileBoundsArgs T{P.keySorted.get(), n, ntiles, C.tileBegin.get(), C.tileEnd.get(), C.tileBound.get(),
                         perCharge};
  kPermute32<<<dim3(blocks, 4), 256, 0, st>>>(A, P.permSorted.get(), n, T);
  launchCheck("permute");
  for (int a = 0; a < 3; ++a) P.cell[a].swap(P.cellT[a]);
  for (int a = 0; a < 9; ++a) P.s[a].swap(P.t[a]);
  P.id.swap(P.idT);
  P.state.swap(P.stateT);
  P.beta.swap(P.betaT);
  const int rz = C.shape == kShapeQuintic ? Box<kShapeQuintic>::RZ : Box<kShapeBinomial3X>::RZ;
  buildChunkLists(C, st, ntx, nty, ntiles, rz);
}

}  // namespace

struct SplitStepContext {
  bool pending = false, middleDone = false, verify = false;
  int n = 0, ntx = 0, nty = 0, ntz = 0, ntiles = 0;
  std::size_t chunkBound = 0;
  unsigned trials = 2, nChunks = 0, occupied = 0, active = 0;
  GArgs32 G{};
  StepArgs S{};
  double o[3] = {0.0, 0.0, 0.0};
};

// Phase 1 of a split step: checks, J clear, sort and chunk lists (their
// totals copied back asynchronously), and the zero-current preview.
// The split positions live on the lattice x = cell * h; the grid origin
// must be one of its nodes: oc its lattice cell, o = oc * h.
static void splitOrigin(const FieldState32& F, const double origin[3], int oc[3], double o[3]) {
  const double h[3] = {F.g.dx, F.g.dy, F.g.dz};
  for (int a = 0; a < 3; ++a) {
    const double r = origin[a] / h[a];
    if (!std::isfinite(r) || std::fabs(r) >= 1.0e9 || std::fabs(r - std::nearbyint(r)) > 1e-6)
      throw std::invalid_argument("An FP32-split grid origin must lie on the cell lattice (origin = integer * spacing)");
    oc[a] = static_cast<int>(std::nearbyint(r));
    o[a] = static_cast<double>(oc[a]) * h[a];
  }
}

// The step's sort phase on stream ss: tile lists and readback words sized,
// error words cleared, then (with particles) the stable tile sort, the
// state permute, the chunk lists and the asynchronous readback of their
// totals and the sort-phase words (C.sorted marks its end).
static void sortPhaseSplit(CouplingState& C, SplitStore& P, const double origin[3], hipStream_t ss) {
  FieldState32& F = *C.fields32;
  int oc[3];
  double o[3];
  splitOrigin(F, origin, oc, o);
  const int n = static_cast<int>(P.count);
  const int ntx = static_cast<int>((F.g.nx + kTileX - 1) / kTileX);
  const int nty = static_cast<int>((F.g.ny + kTileY - 1) / kTileY);
  const int ntz = static_cast<int>((F.g.nz + kTileZ - 1) / kTileZ);
  const int ntiles = ntx * nty * ntz;
  reserveTileLists(C, ss, ntiles, n);
  reserveCouplingWords(C, ss);
  unsigned* pin = static_cast<unsigned*>(C.pinned.get(1024));
  fillInts(ss, reinterpret_cast<int*>(C.error.get()), kErrorWords, 0);
  if (!n) return;
  SortArgs K{};
  K.window = F.window ? 1 : 0;
  K.nz = static_cast<int>(F.g.nz);
  K.lowZ = o[2];
  K.upZ = std::fma(static_cast<double>(F.g.nz), F.g.dz, o[2]);
  K.dz = F.g.dz;
  K.words = C.error.get();
  K.removed = P.removed.get();
  sortSplit(C, P, gatherArgs32(C, F, oc), K, ntx, nty, ntiles, ss);
  hipCheck(hipMemcpyAsync(pin, C.tileScan.get() + ntiles, sizeof(unsigned long long), hipMemcpyDeviceToHost, ss),
           "hipMemcpyAsync");
  hipCheck(hipMemcpyAsync(pin + 120, C.error.get(), 5 * sizeof(unsigned), hipMemcpyDeviceToHost, ss),
           "hipMemcpyAsync");
  hipCheck(hipEventRecord(C.sorted.get(), ss), "hipEventRecord");
}

void prepareStepSplit(CouplingState& C, SplitStore& P, const double origin[3]) {
  FieldState32& F = *C.fields32;
  if (C.prepared.valid) throw std::logic_error("A prepared coupling step is already pending");
  if (C.split && C.split->pending) throw std::logic_error("The previous FP32-split coupling step was not finished");
  hipStream_t ss = C.sortStream.get();
  // J of the last deposit is cleared through that step's active-tile list,
  // which this sort rewrites: clear it now, ahead of the sort (every step
  // starts from J = 0, and the step's own clear then finds nothing to do)
  if (F.footprint.valid && !F.currentZero) launchClearCurrent(F);
  hipCheck(hipEventRecord(C.prepareStart.get(), F.stream), "hipEventRecord");
  hipCheck(hipStreamWaitEvent(ss, C.prepareStart.get(), 0), "hipStreamWaitEvent");
  sortPhaseSplit(C, P, origin, ss);
  hipCheck(hipEventRecord(C.sortReady.get(), ss), "hipEventRecord");
  auto& R = C.prepared;
  R.valid = true;
  R.store = &P;
  R.generation = P.generation;
  R.count = P.count;
  for (int a = 0; a < 3; ++a) R.origin[a] = origin[a];
  R.seq = 0;
}

void splitStepBegin(CouplingState& C, SplitStore& P, const double origin[3], double time, unsigned substeps) {
  FieldState32& F = *C.fields32;
  if (!C.split) C.split = std::make_shared<SplitStepContext>();
  SplitStepContext& X = *C.split;
  if (X.pending) throw std::logic_error("The previous FP32-split coupling step was not finished");
  if (F.window && F.phase != 2) throw std::logic_error("Moving-window electric phase mismatch");
  if (F.window && !F.inletSet) throw std::logic_error("Moving-window inlet planes were not supplied");
  if (substeps == 0) throw std::invalid_argument("Particle substeps must be positive");
  int oc[3];
  double o[3];
  splitOrigin(F, origin, oc, o);
  const int n = static_cast<int>(P.count);
  hipStream_t st = F.stream;
  launchClearCurrent(F);
  for (int a = 0; a < 3; ++a)
    if (C.estar32[a].size() != F.compSize(a)) {
      C.estar32[a].resize(F.compSize(a));
      C.esum32[a].resize(F.compSize(a));
    }
  const GArgs32 G = gatherArgs32(C, F, oc);
  const int ntx = static_cast<int>((F.g.nx + kTileX - 1) / kTileX);
  const int nty = static_cast<int>((F.g.ny + kTileY - 1) / kTileY);
  const int ntz = static_cast<int>((F.g.nz + kTileZ - 1) / kTileZ);
  const int ntiles = ntx * nty * ntz;
  const std::size_t chunkBound = chunkListBound(ntiles, n);
  const unsigned trials = std::max(2u, C.splitTrials);
  if (preparedFor(C, &P, P.generation, P.count, origin)) {
    hipCheck(hipStreamWaitEvent(st, C.sortReady.get(), 0), "hipStreamWaitEvent");
  } else {
    if (C.prepared.valid) hipCheck(hipStreamWaitEvent(st, C.sortReady.get(), 0), "hipStreamWaitEvent");
    sortPhaseSplit(C, P, origin, st);
  }
  C.prepared.valid = false;
  // E* and E* + E^n on the nodes of active tiles only (their only readers)
  float* const out[3] = {C.estar32[0].get(), C.estar32[1].get(), C.estar32[2].get()};
  float* const sums[3] = {C.esum32[0].get(), C.esum32[1].get(), C.esum32[2].get()};
  if (n)
    launchElectricPreviewTiles(F, F.ptrs(), out, sums, F.window ? F.inletPlane[0] : nullptr,
                               F.window ? F.inletPlane[1] : nullptr, C.activeTiles.get(), C.tileScan.get() + ntiles,
                               TileGrid{ntx, nty, ntz, kTileX, kTileY, kTileZ});
  X.pending = true;
  X.middleDone = false;
  X.n = n;
  X.ntx = ntx;
  X.nty = nty;
  X.ntz = ntz;
  X.ntiles = ntiles;
  X.chunkBound = chunkBound;
  X.trials = trials;
  X.G = G;
  for (int a = 0; a < 3; ++a) X.o[a] = o[a];
  X.S = stepArgsFor(C, F.g, F.hasCpml ? &F.cpml : nullptr, F.window, o, time, substeps);
}

// Phase 2: wait for the chunk totals, enqueue the Picard trials and the
// commit, and the one asynchronous readback of the step; the store's
// state arrays are swapped (stream order keeps later work consistent).
void splitStepMiddle(CouplingState& C, SplitStore& P) {
  FieldState32& F = *C.fields32;
  if (!C.split || !C.split->pending || C.split->middleDone) throw std::logic_error("FP32-split step phase mismatch");
  SplitStepContext& X = *C.split;
  int n = X.n;
  const int ntx = X.ntx, nty = X.nty, ntz = X.ntz, ntiles = X.ntiles;
  const std::size_t chunkBound = X.chunkBound;
  unsigned trials = X.trials;
  const GArgs32& G = X.G;
  const double* o = X.o;
  hipStream_t st = F.stream;
  unsigned* pin = static_cast<unsigned*>(C.pinned.get(1024));
  double* pinD = reinterpret_cast<double*>(static_cast<unsigned char*>(C.pinned.get(1024)) + 512);
  (void)ntiles;
  unsigned nChunks = 0, nActive = 0, occupiedTiles = 0;
  if (n) {
    hipCheck(hipEventSynchronize(C.sorted.get()), "hipEventSynchronize");
    unsigned long long total;
    std::memcpy(&total, pin, sizeof(total));
    nChunks = static_cast<unsigned>(chunkField(total));
    nActive = static_cast<unsigned>(activeField(total));
    occupiedTiles = static_cast<unsigned>(occupiedField(total));
    // finishMovingWindowStep's checks of the entries, applied to this origin
    raiseParticleErrors(pin[120]);
    const int behind = static_cast<int>(pin[122]);
    if (pin[123]) cropCounters(C, st, n - behind, n, P.state.get(), P.s[6].get(), P.s[8].get(), true);
    n -= behind;  // removed and cropped entries now trail the live ones
    P.count = static_cast<std::size_t>(n);
    X.n = n;
    // with no active entry (held ones and carriers only) no particle makes a
    // coupling call, and the reference converges in one trial
    // (simulation_source_coupling.cpp:50)
    if (!pin[124]) trials = 1;
    if (nChunks > chunkBound) throw std::logic_error("Chunk list exceeds its bound");
  } else {
    trials = 1;
  }
  const bool quintic = C.shape == kShapeQuintic;
  const int boxN = quintic ? boxNodes<kShapeQuintic>() : boxNodes<kShapeBinomial3X>();
  const std::size_t partialNeed = static_cast<std::size_t>(occupiedTiles) * 3 * boxN;
  reserveStepBuffers(C, partialNeed, nChunks, st);

  const StepArgs& S = X.S;
  PArgs32 A = particleArgs32(P, F);
  A.tileOf = P.keySorted.get();
  A.tileBound = C.tileBound.get();
  const bool verify = C.splitVerify;
  for (unsigned trial = 0; trial < trials; ++trial) {
    unsigned* errT = C.error.get() + kTrialWord + trial;
    const bool last = trial + 1 == trials;
    if (nChunks && trial == 0) {
      ScopedKernel sk(F.profiler, st, "gather");
      if (quintic)
        kGatherChunks32<0, kShapeQuintic><<<nChunks, 2 * kChunk, 0, st>>>(A, G, S, C.chunks.get(), ntx, nty, F.g.dt,
                                                                           nullptr, errT);
      else
        kGatherChunks32<0, kShapeBinomial3X><<<nChunks, 2 * kChunk, 0, st>>>(A, G, S, C.chunks.get(), ntx, nty,
                                                                              F.g.dt, nullptr, errT);
      launchCheck("gather");
    }
    if (n) {
      ScopedKernel sk(F.profiler, st, "push");
      kPushSplit<<<ceilDiv(n, 128), 128, 0, st>>>(A, S, o[0], o[1], o[2], errT);
      launchCheck("push");
    }
    if (!nChunks) {
      referenceCurrentTrial(C, F, C.esum32[2].get(), C.estar32[2].get(), false, false, ntx, nty, ntz);
      continue;
    }
    {
      ScopedKernel sk(F.profiler, st, "deposit");
      if (quintic)
        kDepositSplit<kShapeQuintic><<<nChunks, kChunk * depositLanes<kShapeQuintic>(), 0, st>>>(A, G, S, C.chunks.get(), C.tileBegin.get(),
                                                                 C.tileEnd.get(), C.tileScan.get(), ntx, nty,
                                                                 C.partial.get(), C.tileShift.get(), errT);
      else
        kDepositSplit<kShapeBinomial3X><<<nChunks, kChunk * depositLanes<kShapeBinomial3X>(), 0, st>>>(A, G, S, C.chunks.get(), C.tileBegin.get(),
                                                                    C.tileEnd.get(), C.tileScan.get(), ntx, nty,
                                                                    C.partial.get(), C.tileShift.get(), errT);
      launchCheck("deposit");
    }
    // the last trial's E_cand + E_old is read only by a verifying residual
    const bool esum = !last || verify;
    {
      ScopedKernel sk(F.profiler, st, "reduce_current");
      float* j[3] = {F.base(6), F.base(7), F.base(8)};
#define APRL_REDUCE32(SH, W)                                                                                      \
  kReduceCurrentSplit<SH, W><<<nActive, 256, 0, st>>>(j[0], j[1], j[2], G, C.activeTiles.get(), C.tileScan.get(), \
                                                       C.tileShift.get(), C.partial.get(), ntx, nty, ntz)
      if (quintic && esum) APRL_REDUCE32(kShapeQuintic, true);
      else if (quintic) APRL_REDUCE32(kShapeQuintic, false);
      else if (esum) APRL_REDUCE32(kShapeBinomial3X, true);
      else APRL_REDUCE32(kShapeBinomial3X, false);
#undef APRL_REDUCE32
      launchCheck("reduce current");
    }
    F.currentZero = false;
    F.footprint.valid = true;
    F.footprint.active = C.activeTiles.get();
    F.footprint.count = static_cast<int>(nActive);
    F.footprint.ntx = ntx;
    F.footprint.nty = nty;
    F.footprint.ntz = ntz;
    F.footprint.tx = kTileX;
    F.footprint.ty = kTileY;
    F.footprint.tz = kTileZ;
    // TransverseReferenceCurrent::apply after the trial's deposit
    referenceCurrentTrial(C, F, C.esum32[2].get(), C.estar32[2].get(), esum, true, ntx, nty, ntz);
    if (last && !verify) break;
    {
      ScopedKernel sk(F.profiler, st, "residual");
      double* stats = last ? C.reduceBlocks.get() : nullptr;
      if (quintic)
        kGatherChunks32<1, kShapeQuintic><<<nChunks, 2 * kChunk, 0, st>>>(A, G, S, C.chunks.get(), ntx, nty, F.g.dt,
                                                                           stats, errT);
      else
        kGatherChunks32<1, kShapeBinomial3X><<<nChunks, 2 * kChunk, 0, st>>>(A, G, S, C.chunks.get(), ntx, nty,
                                                                              F.g.dt, stats, errT);
      launchCheck("residual");
    }
    if (last) reduceChunkStats(st, C.reduceBlocks.get(), static_cast<int>(nChunks), C.reduceOut.get());
    for (int a = 0; a < 6; ++a) std::swap(A.fin[a], A.fout[a]);
  }
  // one readback per step: the trial error words (and the verifying
  // residual), final once the last trial has run; the step's end waits for
  // them only, so the commit below runs while the host enqueues the next step
  hipCheck(hipMemcpyAsync(pin + 4, C.error.get() + kTrialWord, trials * sizeof(unsigned), hipMemcpyDeviceToHost, st),
           "hipMemcpyAsync");
  if (verify && nChunks)
    hipCheck(hipMemcpyAsync(pinD, C.reduceOut.get(), 4 * sizeof(double), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
  hipCheck(hipEventRecord(C.stepDone.get(), st), "hipEventRecord");
  // commitConservativeElectric with the last trial's current
  float* const none[3] = {nullptr, nullptr, nullptr};
  if (F.window) windowCommitElectric(F, F.ptrs());
  else launchElectric(F, F.ptrs(), 0, none, F.inletPlane[0], F.inletPlane[1]);
  if (F.window) F.phase = 3;
  X.verify = verify;
  X.trials = trials;
  X.nChunks = nChunks;
  X.occupied = occupiedTiles;
  X.active = nActive;
  X.middleDone = true;
  for (int a = 0; a < 3; ++a) P.cell[a].swap(P.endCell[a]);
  for (int a = 0; a < 6; ++a) P.s[a].swap(P.end[a]);
  C.lastStep.valid = true;
  C.lastStep.store = &P;
  C.lastStep.generation = P.generation;
  for (int a = 0; a < 3; ++a) C.lastStep.origin[a] = o[a];
}

// Phase 3: wait for the step, raise the first device error in trial order,
// fill the report.
void splitStepEnd(CouplingState& C, SplitStore& P, aprl_step_report* report) {
  if (!C.split || !C.split->pending || !C.split->middleDone) throw std::logic_error("FP32-split step phase mismatch");
  SplitStepContext& X = *C.split;
  X.pending = false;
  const unsigned trials = X.trials;
  const bool verify = X.verify;
  const unsigned nChunks = X.nChunks;
  unsigned* pin = static_cast<unsigned*>(C.pinned.get(1024));
  double* pinD = reinterpret_cast<double*>(static_cast<unsigned char*>(C.pinned.get(1024)) + 512);
  hipCheck(hipEventSynchronize(C.stepDone.get()), "hipEventSynchronize");
  for (unsigned t = 0; t < trials; ++t) raiseParticleErrors(pin[4 + t]);
  // Boundary events of the last trial: carriers, removals, counters. The
  // commit only touches entry states, weights and counters, which nothing
  // reads before the next step or query.
  if (X.n && (pin[4 + trials - 1] & kEvtAny))
    commitEvents(C, C.fields32->stream, X.n, P.evt.get(), P.state.get(), P.s[8].get(), P.s[6].get(), P.wend.get(),
                 P.removed.get(), C.error.get() + kTrialWord + (trials - 1), true);
  double force = 0., work = 0.;
  if (verify && nChunks) {
    force = pinD[0];
    work = pinD[3] == 0. ? 0. : (pinD[2] > 0. ? pinD[1] / pinD[2] : pinD[1]);
  }
  if (report) {
    report->iterations = trials;
    report->force = force;
    report->work = work;
    report->tiles = X.occupied;
    report->chunks = nChunks;
    report->active_tiles = X.active;
  }
}

void couplingStepSplit(CouplingState& C, SplitStore& P, const double origin[3], double time, unsigned substeps,
                       aprl_step_report* report) {
  splitStepBegin(C, P, origin, time, substeps);
  splitStepMiddle(C, P);
  splitStepEnd(C, P, report);
}

}  // namespace aprl_gpu
