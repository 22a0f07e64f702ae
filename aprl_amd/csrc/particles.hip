// gfx950 particle kernels of the conservative-source step (FP64 form):
// tile sort, orbit gather + Lapenta-Markidis push with prescribed magnets,
// deterministic tile deposition of the binomial3-x or compensated-quintic
// current, residual gathers with fixed-order reductions, and the Picard
// orchestration.
//
// Per-particle arithmetic follows source_coupling_math.h / source_coupling.cpp
// / longitudinal_shape.h in the reference's association order (contraction
// off), so with identical field inputs a push is bitwise the reference's.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "diagnostics.hpp"
#include "particle_device.hpp"
#include "shapes.hpp"
#include "tuner.hpp"

namespace aprl_gpu {

// ---------------------------------------------------------------------------
// Kernel argument blocks.

struct GArgs {
  double* esum[3];         // E_cand + E_old of the trial
  const double* estar[3];  // zero-current preview E*
  const double* eold[3];   // E^n
  const double* j[3];      // current of the trial
  const double* b[3];      // B^(n+1/2)
  int ne[3][3], nb[3][3];  // (nx, ny, nz) of E (= J) and B components
  int nx, ny, nz;          // cells
  double o[3], h[3];
  double negS;             // (-1*dt)/eps0, applyCurrentSource(-1)
  int smoothX;
  // the Picard stop word: a trial's kernels return at once when it is set
  // (null for work that always runs)
  const unsigned* stop;
};

struct PArgs {
  double* x[3];
  double* u[3];
  const double* q;
  const double* m;
  double* fin[6];   // fields the trial pushes with: Ex Ey Ez Bx By Bz
  double* fout[6];  // next fields of the residual gather
  double* x1[3];
  double* u1[3];
  double* gridAbs;
  const int* state;  // entry states (kStateActive / kStateCarrier for live entries)
  int* evt;          // boundary event code of the trial
  double* cut[3];    // t1, t2, carrier-part deposit weight
  double* wend;      // carrier end weight
  const double* w;   // weight slot (a carrier's current weight)
  const double* beta;  // held speed of held entries
  const unsigned* tileOf;         // sorted tile key of each entry
  unsigned long long* tileBound;  // per tile: largest particle current bound
  double scaleDen;   // dx*dy*dz*dt
  int n;
};

// ---------------------------------------------------------------------------
// Device math restated from the reference.

// clipOpenOrbit for one owner of the whole domain, source_coupling_math.h:235-250.
struct Clip { double b[3], e[3], first, last; bool valid; };
__device__ Clip clipOrbit(const GArgs& G, const double begin[3], const double end[3], unsigned& err) {
  Clip r;
  r.first = 0.;
  r.last = 1.;
  r.valid = true;
  const int size[3] = {G.nx, G.ny, G.nz};
  for (int a = 0; a < 3; ++a) {
    r.b[a] = begin[a];
    r.e[a] = end[a];
    const double x = (begin[a] - G.o[a]) / G.h[a], d = (end[a] - begin[a]) / G.h[a];
    if (!isfinite(x) || !isfinite(d)) { err |= kErrNonfinite; r.valid = false; return r; }
    if (d == 0.) {
      if (x < 0. || x > size[a]) r.valid = false;
      continue;
    }
    {
      const double t = -x / d;
      if (d > 0.) r.first = fmax(r.first, t);
      else r.last = fmin(r.last, t);
    }
    {
      const double t = (double(size[a]) - x) / d;
      if (d > 0.) r.last = fmin(r.last, t);
      else r.first = fmax(r.first, t);
    }
  }
  r.valid = r.valid && r.last >= r.first;
  if (r.valid)
    for (int a = 0; a < 3; ++a) {
      const double d = end[a] - begin[a];
      if (r.first != 0.) r.b[a] = begin[a] + r.first * d;
      if (r.last != 1.) r.e[a] = begin[a] + r.last * d;
    }
  return r;
}

// Impl::electric wide branch (source_coupling.cpp:74-82): orbit average of
// 1/2 (E_cand + E_old) over the clipped chord with the exact adjoint weights
// of the deposit, read from the staged box es.
template <int SHAPE>
__device__ void gatherE(const GArgs& G, const double* es, const TileBox<SHAPE>& box, const double a[3],
                        const double bpt[3], double res[3], unsigned& err) {
  constexpr int N = boxNodes<SHAPE>(), NX = Box<SHAPE>::NX, NY = Box<SHAPE>::NY;
  res[0] = res[1] = res[2] = 0.0;
  const Clip path = clipOrbit(G, a, bpt, err);
  if (!path.valid) { err |= kErrOrbitOutside; return; }
  double x[3], v[3], cuts[kMaxCuts];
  for (int d = 0; d < 3; ++d) {
    x[d] = (path.b[d] - G.o[d]) / G.h[d];
    v[d] = (path.e[d] - path.b[d]) / G.h[d];
  }
  const int count = orbitCuts(x, v, cuts, err);
  const int org[3] = {0, 0, 0};  // x is grid-relative
  int base = 0;
  // all taps in range of every component and inside the box
  auto interior = [&](long long i, long long j, long long k, int wx, int wy, int wz) {
    if (i < 0 || j < 0 || k < 0 || i + wx > G.nx + 1 || j + wy > G.ny + 1 || k + wz > G.nz + 1) return false;
    if (SHAPE == kShapeBinomial3X) {
      // Ex spans x [i, i+3) in nx, Ey y [j, j+1) in ny, Ez z [k, k+3) in nz
      if (i + 3 > G.nx || j + 1 > G.ny || k + 3 > G.nz) return false;
    } else {
      // Ex spans x [i, i+1) in nx, Ey y [j, j+1) in ny, Ez z [k, k+11) in nz
      if (i + 1 > G.nx || j + 1 > G.ny || k + 11 > G.nz) return false;
    }
    if (!box.holds(i, j, k, wx, wy, wz)) return false;
    base = box.base(i, j, k);
    return true;
  };
  auto fast = [&](int d, int oi, int oj, int ok, double w) {
    res[d] += .5 * w * es[d * N + base + (ok * NY + oj) * NX + oi];
  };
  auto visit = [&](int d, long long ii, long long jj, long long kk, double w) {
    if (ii < 0 || jj < 0 || ii >= G.ne[d][0] || jj >= G.ne[d][1]) return;
    if (kk < 0 || kk >= G.ne[d][2]) return;
    const int l = box.local(static_cast<int>(ii), static_cast<int>(jj), static_cast<int>(kk), err);
    res[d] += .5 * w * (l < 0 ? 0.0 : es[d * N + l]);
  };
  for (int s = 1; s < count; ++s) {
    if (cuts[s] - cuts[s - 1] <= 0.) continue;
    segmentTaps<SHAPE, double>(G.nx, G.ny, org, x, v, cuts[s - 1], cuts[s], interior, fast, visit);
  }
  const double f = path.last - path.first;
  res[0] *= f;
  res[1] *= f;
  res[2] *= f;
}

__device__ __forceinline__ int clampIndex(long long v, int n) {
  return static_cast<int>(v < 0 ? 0 : (v > n - 1 ? n - 1 : v));
}

// Impl::magnetic for binomial3-x: binomialXZMagneticSample at the clipped
// chord midpoint with clamped ("open") indices (source_coupling.cpp:83-90;
// source_coupling_math.h:142-167; NativeHalo::openMagneticValue), read
// straight from B in HBM through the caches (the particles of a chunk
// share most of their 4x2x4 stencils), so that the gather stages only E.
__device__ void gatherBBinomial(const GArgs& G, const double a[3], const double bpt[3], double res[3],
                                      unsigned& err) {
  res[0] = res[1] = res[2] = 0.0;
  const Clip path = clipOrbit(G, a, bpt, err);
  if (!path.valid) { err |= kErrMagneticOutside; return; }
  double mid[3];
  for (int d = 0; d < 3; ++d) { mid[d] = path.b[d]; mid[d] += path.e[d]; mid[d] *= .5; }
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    long long cell[3];
    double frac[3];
    for (int d = 0; d < 3; ++d) {
      const double pos = (mid[d] - G.o[d]) / G.h[d] - (comp == d ? 0. : 0.5);
      cell[d] = static_cast<long long>(floor(pos));
      frac[d] = pos - cell[d];
    }
    double wx[4], wz[4];
    nodeWeights(frac[0], wx);
    nodeWeights(frac[2], wz);
    const double wy[2] = {1. - frac[1], frac[1]};
    const int n0 = G.nb[comp][0], n1 = G.nb[comp][1], n2 = G.nb[comp][2];
    const double* bc = G.b[comp];
    if (cell[0] >= 1 && cell[0] + 2 <= n0 - 1 && cell[1] >= 0 && cell[1] + 1 <= n1 - 1 && cell[2] >= 1 &&
        cell[2] + 2 <= n2 - 1) {
      const double* base = bc + (static_cast<size_t>(cell[2] - 1) * n1 + cell[1]) * n0 + (cell[0] - 1);
      for (int k = 0; k < 4; ++k)
        for (int j = 0; j < 2; ++j) {
          const double* r = base + (static_cast<size_t>(k) * n1 + j) * n0;
          double row = 0.;
          for (int i = 0; i < 4; ++i) row += wx[i] * r[i];
          res[comp] += wz[k] * wy[j] * row;
        }
      continue;
    }
    for (int k = 0; k < 4; ++k)
      for (int j = 0; j < 2; ++j) {
        const int kk = clampIndex(cell[2] - 1 + k, n2), jj = clampIndex(cell[1] + j, n1);
        const double* r = bc + (static_cast<size_t>(kk) * n1 + jj) * n0;
        double row = 0.;
        for (int i = 0; i < 4; ++i) row += wx[i] * r[clampIndex(cell[0] - 1 + i, n0)];
        res[comp] += wz[k] * wy[j] * row;
      }
  }
}

// Impl::magnetic for the quintic shape: magneticSample at the clipped chord
// midpoint with clamped indices (source_coupling.cpp:83-86;
// source_coupling_math.h:122-138), read from HBM.
__device__ void gatherBQuintic(const GArgs& G, const double a[3], const double bpt[3], double res[3], unsigned& err) {
  res[0] = res[1] = res[2] = 0.0;
  const Clip path = clipOrbit(G, a, bpt, err);
  if (!path.valid) { err |= kErrMagneticNoDomain; return; }
  double mid[3];
  for (int d = 0; d < 3; ++d) { mid[d] = path.b[d]; mid[d] += path.e[d]; mid[d] *= .5; }
  int cell[2][3];
  double fraction[2][3];
  for (int half = 0; half < 2; ++half)
    for (int d = 0; d < 3; ++d) {
      const double x = (mid[d] - G.o[d]) / G.h[d] - .5 * half;
      cell[half][d] = static_cast<int>(floor(x));
      fraction[half][d] = x - cell[half][d];
    }
#pragma unroll
  for (int axis = 0; axis < 3; ++axis) {
    const int hx = axis == 0 ? 0 : 1, hy = axis == 1 ? 0 : 1, hz = axis == 2 ? 0 : 1;
    const int* n = G.nb[axis];
    const double* src = G.b[axis];
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) {
          const double xw = i ? fraction[hx][0] : 1 - fraction[hx][0];
          const double yw = j ? fraction[hy][1] : 1 - fraction[hy][1];
          const double zw = k ? fraction[hz][2] : 1 - fraction[hz][2];
          const int ii = clampIndex(cell[hx][0] + i, n[0]), jj = clampIndex(cell[hy][1] + j, n[1]),
                    kk = clampIndex(cell[hz][2] + k, n[2]);
          res[axis] += xw * yw * zw * src[(static_cast<size_t>(kk) * n[1] + jj) * n[0] + ii];
        }
  }
}

// ---------------------------------------------------------------------------
// Sort and chunk list.

__device__ __forceinline__ int clampCell(double lattice, int n) {
  const int c = static_cast<int>(floor(lattice));
  return c < 0 ? 0 : (c > n - 1 ? n - 1 : c);
}

__global__ void kTileKeys(const double* x, const double* y, const double* z, int* state, int n, GArgs G, int ntx,
                          int nty, unsigned* key, unsigned* perm, TileRangeArgs R, SortArgs K) {
  __shared__ unsigned blockActive;
  if (threadIdx.x == 0) blockActive = 0u;
  __syncthreads();
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  clearTileRanges(R, p);
  if (p == 0) *K.removed = 0u;
  bool active = false;
  if (p < n) {
    perm[p] = p;
    if (!liveEntry(K, state, p, z[p])) {
      key[p] = static_cast<unsigned>(R.ntiles);
    } else {
      const int cx = clampCell((x[p] - G.o[0]) / G.h[0], G.nx);
      const int cy = clampCell((y[p] - G.o[1]) / G.h[1], G.ny);
      const int cz = clampCell((z[p] - G.o[2]) / G.h[2], G.nz);
      key[p] = (static_cast<unsigned>(cz / kTileZ) * nty + cy / kTileY) * ntx + cx / kTileX;
      active = state[p] == kStateActive;
    }
  }
  // live active entries: wave ballots summed in LDS, one global atomic per
  // block (a global atomic per wave would serialize on the one word)
  const unsigned long long act = __ballot(active);
  if (act && static_cast<int>(__lane_id()) == __ffsll(static_cast<long long>(act)) - 1)
    atomicAdd(&blockActive, static_cast<unsigned>(__popcll(act)));
  __syncthreads();
  if (threadIdx.x == 0 && blockActive) atomicAdd(K.words + 4, blockActive);
}

struct PermuteArgs {
  const double* src[9];
  double* dst[9];
  const uint64_t* idSrc;
  uint64_t* idDst;
  const int* stateSrc;
  int* stateDst;
  const double* betaSrc;
  double* betaDst;
};

// State scatter into sorted order, and the tile ranges of the sorted keys.
// The gathered reads of a permute are latency-bound: blockIdx.y splits the
// arrays over four times as many waves (y = 0: positions, tile ranges and
// bounds; 1: momenta; 2: q, m, w; 3: ids, states, held speeds).
__global__ void kPermute(PermuteArgs A, const unsigned* perm, int n, TileBoundsArgs T) {
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
      tileBoundsAt(T, p);
      const unsigned k = T.key[p];
      if (k < static_cast<unsigned>(T.ntiles)) {
        tile = k;
        bound = entryCurrentBound(A.src[6][from], A.stateSrc[from], T.perCharge);
      }
    }
  }
  if (part == 0) raiseTileBound(T.bound, tile, bound);
}

// Per tile, packed (packTileCounts): chunk count, whether it is occupied,
// and whether any tile within one tile in x and y and rz tiles in z is
// occupied (its J nodes can then receive current); entry ntiles is 0 so
// that the exclusive scan ends with the totals.
__global__ void kTileCounts(const int* begin, const int* end, int ntx, int nty, int ntz, int rz,
                            unsigned long long* counts) {
  const int ntiles = ntx * nty * ntz;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  if (t == ntiles) {
    counts[t] = 0ull;
    return;
  }
  const int tx = t % ntx, ty = (t / ntx) % nty, tz = t / (ntx * nty);
  unsigned touched = 0;
  for (int oz = -rz; oz <= rz; ++oz)
    for (int oy = -1; oy <= 1; ++oy)
      for (int ox = -1; ox <= 1; ++ox) {
        const int x = tx + ox, y = ty + oy, z = tz + oz;
        if (x < 0 || x >= ntx || y < 0 || y >= nty || z < 0 || z >= ntz) continue;
        const int s = (z * nty + y) * ntx + x;
        if (end[s] > begin[s]) touched = 1;
      }
  const unsigned chunks = static_cast<unsigned>((end[t] - begin[t] + kChunk - 1) / kChunk);
  counts[t] = packTileCounts(chunks, chunks ? 1u : 0u, touched);
}

// Chunk list in tile order and the list of active tiles (tiles with an
// occupied neighbour), both from the scanned counts.
// With a mailbox, lane 0 also posts the scan total (chunks, active tiles,
// occupied tiles) and the sort-phase words to host memory (values at words
// 8..14, sequence number in flag word 0).
__global__ void kChunkList(const int* begin, const int* end, const unsigned long long* counts,
                           const unsigned long long* scan, int ntiles, int4* chunks, int* active, unsigned* box,
                           unsigned seq, const unsigned* words) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (box && t == 0) {
    const unsigned long long total = scan[ntiles];
    volatile unsigned* v = box;
    v[8] = static_cast<unsigned>(chunkField(total));
    v[9] = static_cast<unsigned>(activeField(total));
    for (int w = 0; w < 5; ++w) v[10 + w] = words[w];
    v[11] = static_cast<unsigned>(occupiedField(total));  // in place of word 1
    __threadfence_system();
    v[0] = seq;
  }
  if (t >= ntiles) return;
  const unsigned long long sc = scan[t];
  if (activeField(counts[t])) active[activeField(sc)] = t;
  const int e = end[t];
  int c = chunkField(sc);
  for (int p = begin[t]; p < e; p += kChunk, ++c) chunks[c] = make_int4(t, p, min(e, p + kChunk), 0);
}

__global__ void kFillInt(int* p, int n, int v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

// ---------------------------------------------------------------------------
// Gathers, push, deposit, residual.

// One Picard trial's residual: the chunk partials reduced in a fixed tree
// to (max force, defect sum, scale sum, coupling calls) in out. With a
// Picard word (W.stop) thread 0 then takes the reference's convergence
// decision on the device (simulation_source_coupling.cpp:46-51): at least
// two trials when a particle made a coupling call, force < 1e-13 and
// work < 1e-8, the same double arithmetic as the host test; a trial whose
// error word holds a device error also stops the trials. Once the word is
// set, later trials skip the reduction. With a mailbox, thread 0 posts the
// sums (doubles at byte 2048), the stop word (word 15) and the trial words
// firstWord..W.trial (words 16..) to host memory, sequence number in flag
// word 1, whether or not the trial ran. Called by all 256 threads of one
// workgroup; sh holds 4 x 256 doubles.
struct PicardWords {
  unsigned* stop;
  unsigned trial;
  const unsigned* trialWords;  // error words of trials 0..kMaxTrials-1
  unsigned firstWord;
};

__device__ void residualTail(const double* blockIn, int blocks, double* out, unsigned* box, unsigned seq,
                             const PicardWords& W, double* sh) {
  double* sf = sh;
  double* sd = sh + 256;
  double* ss = sh + 512;
  double* sc = sh + 768;
  const bool skip = W.stop && *W.stop;
  if (!skip) {
    double f = 0., d = 0., s = 0., c = 0.;
    for (int b = threadIdx.x; b < blocks; b += 256) {
      f = fmax(f, blockIn[4 * b]);
      d += blockIn[4 * b + 1];
      s += blockIn[4 * b + 2];
      c += blockIn[4 * b + 3];
    }
    sf[threadIdx.x] = f; sd[threadIdx.x] = d; ss[threadIdx.x] = s; sc[threadIdx.x] = c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w) {
        sf[threadIdx.x] = fmax(sf[threadIdx.x], sf[threadIdx.x + w]);
        sd[threadIdx.x] += sd[threadIdx.x + w];
        ss[threadIdx.x] += ss[threadIdx.x + w];
        sc[threadIdx.x] += sc[threadIdx.x + w];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      out[0] = sf[0]; out[1] = sd[0]; out[2] = ss[0]; out[3] = sc[0];
      if (W.stop) {
        const double force = sf[0];
        const double work = sc[0] == 0. ? 0. : (ss[0] > 0. ? sd[0] / ss[0] : sd[0]);
        const unsigned iterations = W.trial + 1;
        if (W.trialWords[W.trial] & ~kEvtAny) *W.stop = iterations | kStopError;
        else if (iterations >= (sc[0] > 0. ? 2u : 1u) && force < 1e-13 && work < 1e-8) *W.stop = iterations;
      }
    }
  }
  if (threadIdx.x == 0 && box) {
    volatile double* vd = reinterpret_cast<volatile double*>(reinterpret_cast<unsigned char*>(box) + 2048);
    vd[0] = out[0]; vd[1] = out[1]; vd[2] = out[2]; vd[3] = out[3];
    volatile unsigned* v = box;
    v[15] = W.stop ? *W.stop : 0u;
    for (unsigned w = W.firstWord; w <= W.trial; ++w) v[16 + w - W.firstWord] = W.trialWords[w];
    __threadfence_system();
    v[1] = seq;
  }
}

// Chunked orbit gathers from LDS-staged tile fields. A workgroup of
// 2*kChunk lanes takes one chunk: the first kChunk lanes (whole waves)
// gather E, the other kChunk lanes gather B, one lane per particle each.
// MODE 0: the first trial's gather at (x0, ballistic guess) against the
// zero-current preview (source_coupling.cpp:133-147); it writes the fields
// the first push uses. MODE 1: the residual gather of localResidual
// (source_coupling.cpp:148-202) at (x0, x1): next fields, and the chunk's
// force max and defect / scale / call sums in a fixed tree order.
// Binomial3-x stages only E (31 KB of LDS): three waves per SIMD fit once
// the residual gather keeps its registers within 168.
template <int MODE, int SHAPE>
__global__ void __launch_bounds__(2 * kChunk) __attribute__((amdgpu_waves_per_eu(SHAPE == kShapeBinomial3X ? 3 : 1)))
kGatherChunks(PArgs P, GArgs G, StepArgs S, const int4* chunks, int ntx, int nty, double dt, double* chunkOut,
              unsigned* errOut) {
  if (G.stop && *G.stop) return;
  __shared__ double es[3 * boxNodes<SHAPE>()];
  __shared__ double bs[4 * kChunk];
  const int cb = xcdContiguous(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
  const int4 ch = chunks[cb];
  const TileBox<SHAPE> box = tileBox<SHAPE>(ch.x, ntx, nty);
  const bool bWaves = threadIdx.x >= kChunk;
  const int lane = static_cast<int>(threadIdx.x) - (bWaves ? kChunk : 0);
  const int p = ch.y + lane;
  // This lane's particle data is loaded before the tile is staged, so that
  // its latency overlaps the staging loads' (lanes past the chunk's end read
  // its first entry and ignore it).
  const int pv = p < ch.z ? p : ch.y;
  const int stored = P.state[pv];
  const double betaP = P.beta[pv];
  double x0[3], uv[3], fin[6] = {0., 0., 0., 0., 0., 0.}, qm[3] = {0., 0., 0.};
  for (int d = 0; d < 3; ++d) {
    x0[d] = P.x[d][pv];
    uv[d] = MODE == 0 ? P.u[d][pv] : P.x1[d][pv];  // u (first gather) or the trial end x1
  }
  if (MODE == 1 && !bWaves) {
    for (int d = 0; d < 6; ++d) fin[d] = P.fin[d][pv];
    qm[0] = P.q[pv];
    qm[1] = P.m[pv];
    qm[2] = P.gridAbs[pv];
  }
  stageTileE<SHAPE>(G, box, es);
  __syncthreads();
  // carriers move without fields and take no part in the residual; a held
  // entry takes part from its release point with the rest of the step
  const int st = p < ch.z ? stored : kStateRemoved;
  bool live = st == kStateActive;
  double dtp = dt;
  unsigned err = 0;
  double a[3] = {0., 0., 0.}, b[3] = {0., 0., 0.}, g[3] = {0., 0., 0.};
  if (st == kStateActive || st == kStateHeld)
    for (int d = 0; d < 3; ++d) a[d] = x0[d];
  if (st == kStateHeld) {
    const double uh[3] = {P.u[0][p], P.u[1][p], P.u[2][p]};
    const HeldMove h = heldMove(S, a, uh, betaP);
    if (h.release && h.remaining > 0.0) {
      live = true;
      for (int d = 0; d < 3; ++d) a[d] = h.x1[d];
      dtp = h.remaining;
    }
  }
  if (live) {
    if (MODE == 0) {
      const double gamma = gammaFromU(uv[0], uv[1], uv[2]);
      const double s = dtp * kC / gamma;
      for (int d = 0; d < 3; ++d) {
        b[d] = a[d];
        b[d] += s * uv[d];
      }
    } else {
      for (int d = 0; d < 3; ++d) b[d] = uv[d];
    }
    if (bWaves) {
      if constexpr (SHAPE == kShapeBinomial3X) gatherBBinomial(G, a, b, g, err);
      else gatherBQuintic(G, a, b, g, err);
    } else {
      gatherE<SHAPE>(G, es, box, a, b, g, err);
    }
    double* const* dst = MODE == 0 ? P.fin : P.fout;
    const int off = bWaves ? 3 : 0;
    for (int d = 0; d < 3; ++d) dst[off + d][p] = g[d];
  }
  if (MODE == 1) {
    // The staged box is no longer read: reuse it to pass B to the E lanes,
    // then for the chunk reduction.
    __syncthreads();
    double* xb = es;  // 3 * kChunk
    if (bWaves)
      for (int d = 0; d < 3; ++d) xb[d * kChunk + lane] = g[d];
    __syncthreads();
    double force = 0., defect = 0., scale = 0., calls = 0.;
    if (!bWaves && live) {
      const double nb[3] = {xb[lane], xb[kChunk + lane], xb[2 * kChunk + lane]};
      double de[3], db[3];
      for (int d = 0; d < 3; ++d) {
        de[d] = g[d] - fin[d];
        db[d] = nb[d] - fin[3 + d];
      }
      const double q = qm[0], m = qm[1];
      force = fabs(q / m) * dtp / kC * (sqrt(de[0] * de[0] + de[1] * de[1] + de[2] * de[2]) +
                                        kC * sqrt(db[0] * db[0] + db[1] * db[1] + db[2] * db[2]));
      double w1 = 0., w2 = 0.;
      for (int d = 0; d < 3; ++d) {
        w1 += q * de[d] * (b[d] - a[d]);
        w2 += q * g[d] * (b[d] - a[d]);
      }
      defect = fabs(w1);
      scale = qm[2] + fabs(w2);
      calls = 1.;
    }
    chunkStatsTree(bs, !bWaves, lane, force, defect, scale, calls, chunkOut + 4 * cb);
  }
  if (err) atomicOr(errOut, err);
}


// Push of source_coupling.cpp:133-147 for one trial with the gathered
// fields (the substeps with prescribed magnets), or the straight move of a
// CPML carrier (position += dt u c / gamma, simulation.cpp:4407-4413), then
// the segment's boundary events. (The tile current bounds the deposit's
// fixed-point exponents come from are set once per step by the sort.)
__global__ void __launch_bounds__(128) kPush(PArgs P, GArgs G, StepArgs S, unsigned* errOut) {
  if (G.stop && *G.stop) return;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  unsigned err = 0;
  const double x0[3] = {P.x[0][p], P.x[1][p], P.x[2][p]};
  const double u0[3] = {P.u[0][p], P.u[1][p], P.u[2][p]};
  double x[3], u[3], absGrid, t1, t2, wDeposit, wEnd;
  const double usedE[3] = {P.fin[0][p], P.fin[1][p], P.fin[2][p]};
  const double usedB[3] = {P.fin[3][p], P.fin[4][p], P.fin[5][p]};
  const int code = entryStep(S, G.o, P.state[p], x0, u0, P.q[p], P.m[p], P.w[p], P.beta[p], usedE, usedB, x, u,
                             absGrid, t1, t2, wDeposit, wEnd, err);
  for (int a = 0; a < 3; ++a) {
    P.x1[a][p] = x[a];
    P.u1[a][p] = u[a];
  }
  P.gridAbs[p] = absGrid;
  P.evt[p] = code;
  if (code) {
    P.cut[0][p] = t1;
    P.cut[1][p] = t2;
    P.cut[2][p] = wDeposit;
    P.wend[p] = wEnd;
    err |= kEvtAny;
  }
  if (err) atomicOr(errOut, err);
}

// Deterministic tile deposition of Current::deposit (longitudinal_shape.h:
// 175-187): one workgroup per chunk, depositLanes<SHAPE>() lanes per
// particle (one tap slice each), every tap of the chord x0 -> x1
// accumulated into an LDS copy of the tile box in 64-bit fixed point. The binary exponent comes from a bound on the whole tile's
// current (the shape's weight bound times the tile's particle count times
// the largest per-particle bound the push recorded), so every chunk of a
// tile uses the same scale. Integer addition is associative, so the LDS
// atomics and the chunk sums are reproducible in any order; each tap is
// rounded once, to 2^-62 of the tile bound.
template <int SHAPE>
__global__ void __launch_bounds__(kChunk * depositLanes<SHAPE>()) __attribute__((amdgpu_waves_per_eu(SHAPE == kShapeBinomial3X ? 3 : 1))) kDeposit(PArgs P, GArgs G, StepArgs S, const int4* chunks,
                                                   const int* tileBegin, const int* tileEnd,
                                                   const unsigned long long* tileScan, int ntx, int nty,
                                                   long long* partial, int* tileShift, unsigned* errOut) {
  if (G.stop && *G.stop) return;
  constexpr int N = boxNodes<SHAPE>(), NX = Box<SHAPE>::NX, NY = Box<SHAPE>::NY;
  __shared__ unsigned long long acc[3 * N];
  const int4 ch = chunks[blockIdx.x];
  const int tile = ch.x;
  const TileBox<SHAPE> box = tileBox<SHAPE>(tile, ntx, nty);
  for (int t = threadIdx.x; t < 3 * N; t += blockDim.x) acc[t] = 0ull;
  const unsigned long long tsc = tileScan[tile];
  const int shift = tileFixedShift(Box<SHAPE>::kWeightBound, P.tileBound[tile], tileEnd[tile] - tileBegin[tile]);
  if (threadIdx.x == 0 && static_cast<int>(blockIdx.x) == chunkField(tsc)) tileShift[tile] = shift;
  __syncthreads();
  unsigned err = 0;
  // lanes [s * kChunk, (s + 1) * kChunk) take tap slice s of particles ch.y + lane
  const int sub = static_cast<int>(threadIdx.x) / kChunk;
  const int p = ch.y + static_cast<int>(threadIdx.x) % kChunk;
  if (p < ch.z) {
    double scale[3];
    double x[3], v[3], cuts[kMaxCuts];
    const int org[3] = {0, 0, 0};  // x is grid-relative
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
    double mult = 1.0;  // weight of the chord part being deposited
    auto fast = [&](int d, int oi, int oj, int ok, double w) {
      if (scale[d] == 0.) return;
      atomicAdd(&acc[d * N + base + (ok * NY + oj) * NX + oi],
                static_cast<unsigned long long>(llrint(ldexp(scale[d] * w * mult, shift))));
    };
    auto visit = [&](int d, long long ii, long long jj, long long kk, double w) {
      if (scale[d] == 0.) return;
      if (ii < 0 || ii >= G.ne[d][0]) {
        if (SHAPE != kShapeBinomial3X || !G.smoothX) err |= kErrCurrentX;
        return;
      }
      if (jj < 0 || jj >= G.ne[d][1]) { err |= kErrCurrentY; return; }
      if (kk < 0 || kk >= G.ne[d][2]) return;
      const int l = box.local(static_cast<int>(ii), static_cast<int>(jj), static_cast<int>(kk), err);
      if (l < 0) return;
      atomicAdd(&acc[d * N + l], static_cast<unsigned long long>(llrint(ldexp(scale[d] * w * mult, shift))));
    };
    // A boundary event splits the last leg: [0, t1] at unit weight (the
    // active part) and [t1, t2] at the carrier part's deposit weight. A held
    // entry released inside the step deposits its held leg x0 -> kink first.
    const int e = P.evt[p];
    const double t1 = e ? P.cut[0][p] : 1.0, t2 = e ? P.cut[1][p] : 1.0, wCarrier = e ? P.cut[2][p] : 0.0;
    double xa[3], xb[3], xk[3];
    for (int d = 0; d < 3; ++d) {
      xa[d] = P.x[d][p];
      xb[d] = P.x1[d][p];
      xk[d] = xa[d];
    }
    int legs = 1;
    if (P.state[p] == kStateHeld) {
      const double uh[3] = {P.u[0][p], P.u[1][p], P.u[2][p]};
      const HeldMove h = heldMove(S, xa, uh, P.beta[p]);
      if (h.release && h.kink) {
        legs = 2;
        for (int d = 0; d < 3; ++d) xk[d] = h.x1[d];
      }
    }
    const double q = P.q[p];
#pragma unroll 1
    for (int leg = 0; leg < legs; ++leg) {
      const bool last = leg + 1 == legs;
      for (int d = 0; d < 3; ++d) {
        const double s0 = leg ? xk[d] : xa[d], s1 = last ? xb[d] : xk[d];
        // scale_a = q * (x1_a - x0_a) / (dx dy dz dt), eb_deposition.cpp:41-51
        scale[d] = q * (s1 - s0) / P.scaleDen;
        x[d] = (s0 - G.o[d]) / G.h[d];
        v[d] = (s1 - s0) / G.h[d];
      }
      const int count = orbitCuts(x, v, cuts, err);
      const double f1 = last ? t1 : 1.0, f2 = last ? t2 : 1.0, fw = last ? wCarrier : 0.0;
      for (int sgm = 1; sgm < count; ++sgm) {
        const double c0 = cuts[sgm - 1], c1 = cuts[sgm];
        if (c1 - c0 <= 0.) continue;
        // without an event: one part [c0, c1] (t1 = t2 = 1)
#pragma unroll 1
        for (int part = 0; part < 2; ++part) {
          const double a = part ? fmax(c0, f1) : c0;
          const double b = part ? fmin(c1, f2) : fmin(c1, f1);
          if (!(b > a)) continue;
          mult = part ? fw : 1.0;
          segmentTaps<SHAPE, double, depositLanes<SHAPE>()>(G.nx, G.ny, org, x, v, a, b, interior, fast, visit, sub);
        }
      }
    }
  }
  __syncthreads();
  flushTileBox<SHAPE>(acc, partial, tileScan, tile);
  if (err) atomicOr(errOut, err);
}

// Node-owner reduction into J, one workgroup per active cell tile (a tile
// with an occupied tile within its box reach): each owned node sums, over
// the covering occupied tiles in fixed (tz, ty, tx) order, the tile's exact
// integer total converted with the tile's exponent, and clears what it read
// for the next deposit. It also writes E_cand + E_old of the trial
// (applyCurrentSource(-1) on the preview, source_coupling.cpp:84-89) for the
// next residual gather. Nodes of other tiles keep the zeros J held at the
// start of the step.
template <int SHAPE>
__global__ void __launch_bounds__(256) kReduceCurrent(double* jx, double* jy, double* jz, GArgs G, const int* active,
                                                      const unsigned long long* tileScan, const int* tileShift,
                                                      long long* partial, int ntx, int nty, int ntz) {
  if (G.stop && *G.stop) return;
  constexpr int RZ = Box<SHAPE>::RZ, NB = 9 * (2 * RZ + 1);
  const int tile = active[xcdContiguous(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x))];
  const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
  __shared__ int nbFirst[NB], nbShift[NB];
  coveringTiles<SHAPE>(tileScan, tileShift, tx, ty, tz, ntx, nty, ntz, nbFirst, nbShift);
  __syncthreads();
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int n0 = G.ne[d][0], n1 = G.ne[d][1], n2 = G.ne[d][2];
    // owned node box (the last tile of an axis also owns the extra node)
    const int i0 = tx * kTileX, i1 = tx == ntx - 1 ? n0 : min(n0, i0 + kTileX);
    const int j0 = ty * kTileY, j1 = ty == nty - 1 ? n1 : min(n1, j0 + kTileY);
    const int k0 = tz * kTileZ, k1 = tz == ntz - 1 ? n2 : min(n2, k0 + kTileZ);
    const int wx = i1 - i0, wy = j1 - j0, wz = k1 - k0;
    double* dst = d == 0 ? jx : (d == 1 ? jy : jz);
    for (int t = threadIdx.x; t < wx * wy * wz; t += blockDim.x) {
      const int i = i0 + t % wx, j = j0 + (t / wx) % wy, k = k0 + t / (wx * wy);
      const double sum = ownedNodeSum<SHAPE>(partial, nbFirst, nbShift, tx, ty, tz, d, i, j, k);
      const size_t idx = (static_cast<size_t>(k) * n1 + j) * n0 + i;
      dst[idx] = sum;
      const bool applied = d == 0 ? (j > 0 && j < G.ny && k != 0 && k != G.nz)
                                  : (d == 1 ? (i > 0 && i < G.nx && k != 0 && k != G.nz)
                                            : (i > 0 && i < G.nx && j > 0 && j < G.ny));
      double e = G.estar[d][idx];
      if (applied && sum != 0.) e += G.negS * sum;
      G.esum[d][idx] = e + G.eold[d][idx];
    }
  }
}

__global__ void __launch_bounds__(256) kReduceBlocks(const double* blockIn, int blocks, double* out, unsigned* box,
                                                     unsigned seq, PicardWords W) {
  __shared__ double sh[4 * 256];
  residualTail(blockIn, blocks, out, box, seq, W, sh);
}

// ---------------------------------------------------------------------------
// Host orchestration.

static GArgs gatherArgs(const CouplingState& C, const FieldState& F, const double origin[3]) {
  GArgs G{};
  for (int a = 0; a < 3; ++a) {
    G.esum[a] = C.esum[a].get();
    G.estar[a] = C.estar[a].get();
    G.eold[a] = F.base(a);
    G.j[a] = F.base(6 + a);
    G.b[a] = F.base(3 + a);
    for (int d = 0; d < 3; ++d) {
      G.ne[a][d] = static_cast<int>(F.dims[a][d]);
      G.nb[a][d] = static_cast<int>(F.dims[3 + a][d]);
    }
    G.o[a] = origin[a];
  }
  G.nx = static_cast<int>(F.g.nx);
  G.ny = static_cast<int>(F.g.ny);
  G.nz = static_cast<int>(F.g.nz);
  G.h[0] = F.g.dx; G.h[1] = F.g.dy; G.h[2] = F.g.dz;
  G.negS = -1.0 * F.g.dt / kEps0;
  G.smoothX = C.smoothX ? 1 : 0;
  return G;
}

static PArgs particleArgs(ParticleStore& P) {
  PArgs A{};
  for (int a = 0; a < 3; ++a) {
    A.x[a] = P.s[a].get();
    A.u[a] = P.s[3 + a].get();
    A.x1[a] = P.end[a].get();
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
  A.n = static_cast<int>(P.count);
  return A;
}

void reserveParticles(ParticleStore& P, std::size_t capacity) {
  if (capacity <= P.capacity) return;
  for (int a = 0; a < 9; ++a) { P.s[a].resize(capacity); P.t[a].resize(capacity); }
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

// Grows the store to `capacity` entries keeping the persistent state of
// its first `count` entries (SoA, ids, entry states, held speeds); the
// per-step arrays are rewritten by every step.
void growParticles(ParticleStore& P, std::size_t capacity) {
  if (capacity <= P.capacity) return;
  DeviceArray<double> s[9], beta;
  DeviceArray<uint64_t> id;
  DeviceArray<int> state;
  for (int a = 0; a < 9; ++a) s[a].swap(P.s[a]);
  id.swap(P.id);
  state.swap(P.state);
  beta.swap(P.beta);
  reserveParticles(P, capacity);
  const std::size_t n = P.count;
  if (!n) return;
  for (int a = 0; a < 9; ++a)
    hipCheck(hipMemcpy(P.s[a].get(), s[a].get(), n * sizeof(double), hipMemcpyDeviceToDevice), "hipMemcpy");
  hipCheck(hipMemcpy(P.id.get(), id.get(), n * sizeof(uint64_t), hipMemcpyDeviceToDevice), "hipMemcpy");
  hipCheck(hipMemcpy(P.state.get(), state.get(), n * sizeof(int), hipMemcpyDeviceToDevice), "hipMemcpy");
  hipCheck(hipMemcpy(P.beta.get(), beta.get(), n * sizeof(double), hipMemcpyDeviceToDevice), "hipMemcpy");
}

void appendParticles(ParticleStore& P, std::size_t k, const double* x3, const double* u3, const double* q,
                     const double* m, const double* w, const uint64_t* id, const int* state, const double* beta) {
  if (!k) return;
  const std::size_t n = P.count;
  if (n + k > P.capacity) growParticles(P, std::max(n + k, P.capacity + P.capacity / 4 + 1024));
  std::vector<double> col(k);
  for (int a = 0; a < 3; ++a) {
    for (std::size_t i = 0; i < k; ++i) col[i] = x3[3 * i + a];
    hipCheck(hipMemcpy(P.s[a].get() + n, col.data(), k * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    for (std::size_t i = 0; i < k; ++i) col[i] = u3[3 * i + a];
    hipCheck(hipMemcpy(P.s[3 + a].get() + n, col.data(), k * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
  }
  const double* scalars[3] = {q, m, w};
  for (int a = 0; a < 3; ++a)
    hipCheck(hipMemcpy(P.s[6 + a].get() + n, scalars[a], k * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
  hipCheck(hipMemcpy(P.id.get() + n, id, k * sizeof(uint64_t), hipMemcpyHostToDevice), "hipMemcpy");
  hipCheck(hipMemcpy(P.state.get() + n, state, k * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
  hipCheck(hipMemcpy(P.beta.get() + n, beta, k * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
  P.count = n + k;
  ++P.generation;
}

void raiseParticleErrors(unsigned e) {
  e &= ~kEvtAny;
  if (!e) return;
  if (e & kErrNonfinite) throw std::runtime_error("Nonfinite open coupling orbit");
  if (e & kErrBehindInlet) throw std::runtime_error("Active particle fell behind moving inlet");
  if (e & kErrCarrierBehind) throw std::runtime_error("Current carrier fell behind moving inlet");
  if (e & kErrRemapOutside)
    throw std::runtime_error("Moving-window remap particle crossed a non-adjacent slab or outer boundary");
  if (e & kErrUpstream)
    throw std::runtime_error("An active particle is upstream of the moving window; the inlet guard did not contain its release event");
  if (e & kErrStartInCpml) throw std::out_of_range("Active particle starts inside the non-physical CPML region");
  if (e & kErrStartOutside) throw std::out_of_range("Particle trajectory starts outside the global E/B domain");
  if (e & kErrBeyondExit) throw std::runtime_error("Active particle started beyond trusted moving-window exit");
  if (e & kErrCpmlLocate) throw std::runtime_error("Cannot locate the first particle entry into CPML");
  if (e & kErrExitLocate) throw std::runtime_error("Cannot locate the first particle-domain exit");
  if (e & kErrEndpointOutside) throw std::out_of_range("Particle trajectory endpoint is outside the E/B domain");
  if (e & kErrOrbitOutside) throw std::out_of_range("Quintic active orbit outside domain");
  if (e & kErrMagneticOutside) throw std::out_of_range("Notch magnetic orbit outside domain");
  if (e & kErrMagneticNoDomain) throw std::out_of_range("Active coupling magnetic orbit has no in-domain portion");
  if (e & kErrRootNonphysical) throw std::runtime_error("Nonphysical local mean-gamma root");
  if (e & kErrRootNoConverge) throw std::runtime_error("Local mean-gamma root did not converge");
  if (e & kErrCurrentY) throw std::out_of_range("Transverse current support outside y grid");
  if (e & kErrCurrentX) throw std::out_of_range("Transverse current support outside x grid");
  if (e & kErrCrossings) throw std::runtime_error("Orbit exceeds the device crossing budget");
  if (e & kErrTileHalo) throw std::runtime_error("Particle chord is longer than one cell per axis in one step");
  throw std::runtime_error("Unknown device error");
}

StepArgs stepArgsFor(const CouplingState& C, const aprl_grid_geometry& g, const aprl_cpml_parameters* cp,
                     bool window, const double origin[3], double time, unsigned substeps) {
  StepArgs S{};
  S.time = time;
  S.dt = g.dt;
  S.substeps = static_cast<int>(substeps);
  S.frame = C.frame;
  S.nMagnets = static_cast<int>(C.magnets.size());
  for (int n = 0; n < S.nMagnets; ++n) S.mag[n] = C.magnets[n];
  const double sp[3] = {g.dx, g.dy, g.dz};
  const std::size_t cells[3] = {g.nx, g.ny, g.nz};
  const double eps = std::numeric_limits<double>::epsilon();
  S.cpmlEnabled = 0;
  S.cpmlOrder = cp ? cp->polynomial_order : 0.0;
  for (int a = 0; a < 3; ++a) {
    const double lo = origin[a], hi = origin[a] + static_cast<double>(cells[a]) * sp[a];
    if (2 * C.cpmlCells[a] >= cells[a] && C.cpmlCells[a] > 0)
      throw std::invalid_argument("Particle CPML region must leave a physical interior");
    S.physLo[a] = lo + static_cast<double>(C.cpmlCells[a]) * sp[a];
    S.physHi[a] = hi - static_cast<double>(C.cpmlCells[a]) * sp[a];
    S.cpmlTol[a] = 64.0 * eps * std::max(1.0, std::max(std::abs(lo), std::abs(hi)));
    S.cpmlOuterLo[a] = lo;
    S.cpmlOuterHi[a] = hi;
    S.cpmlCells[a] = static_cast<int>(C.cpmlCells[a]);
    S.cpmlRate[a] = 0.0;
    S.cpmlThick[a] = static_cast<double>(C.cpmlCells[a]) * sp[a];
    if (C.cpmlCells[a]) {
      // EBCPMLParameters::maximumConductivityRate, eb_cpml.cpp:263-281
      S.cpmlEnabled = 1;
      const double order = cp ? cp->polynomial_order : 0.0, reflection = cp ? cp->target_reflection : 0.0;
      if (!(order > 0.0) || !(reflection > 0.0 && reflection < 1.0))
        throw std::invalid_argument("Cannot evaluate an invalid CPML conductivity profile");
      S.cpmlRate[a] = -(order + 1.0) * kC * std::log(reflection) / (2.0 * S.cpmlThick[a]);
    }
    S.outLo[a] = origin[a];
    S.outHi[a] = std::fma(static_cast<double>(cells[a]), sp[a], origin[a]);
    S.outTol[a] = 64.0 * eps * std::max(1.0, std::max(std::abs(S.outLo[a]), std::abs(S.outHi[a])));
    S.extent[a] = static_cast<double>(cells[a]) * sp[a];
    S.posTol[a] = 64.0 * eps * std::max(1.0, std::max(std::abs(origin[a]), S.extent[a]));
  }
  S.hasExit = C.hasExit ? 1 : 0;
  S.window = window ? 1 : 0;
  S.exitLabZ = C.exitLabZ;
  S.hasActivation = C.hasActivation ? 1 : 0;
  S.activationLabZ = C.activationLabZ;
  S.holdTransverse = C.holdTransverse ? 1 : 0;
  return S;
}

static int bitsFor(unsigned v) {
  int b = 1;
  while (b < 32 && (v >> b) != 0u) ++b;
  return b;
}

// Stable sort of the particle state by start tile (removed and cropped
// entries last), then the tile ranges and the chunk list (tile order,
// kChunk particles per chunk) in device memory.
static void sortAndChunk(CouplingState& C, ParticleStore& P, const GArgs& G, const SortArgs& K, int ntx, int nty,
                         int ntiles, unsigned seq, hipStream_t st) {
  FieldState& F = *C.fields;
  const int n = static_cast<int>(P.count);
  ScopedKernel sk(F.profiler, st, "sort");
  const unsigned blocks = ceilDiv(n, 256);
  kTileKeys<<<blocks, 256, 0, st>>>(
      P.s[0].get(), P.s[1].get(), P.s[2].get(), P.state.get(), n, G, ntx, nty, P.key.get(), P.perm.get(),
      TileRangeArgs{C.tileBegin.get(), C.tileEnd.get(), ntiles, C.tileBound.get()}, K);
  launchCheck("tile keys");
  const int bits = bitsFor(static_cast<unsigned>(ntiles));  // keys 0..ntiles (ntiles: behind the live entries)
  std::size_t bytes = 0;
  tileSort(nullptr, bytes, P.key.get(), P.keySorted.get(), P.perm.get(), P.permSorted.get(), static_cast<unsigned>(n), bits, st);
  if (bytes > P.sortTempBytes) {
    P.sortTemp.resize(bytes);
    P.sortTempBytes = bytes;
  }
  bytes = P.sortTempBytes;
  tileSort(P.sortTemp.get(), bytes, P.key.get(), P.keySorted.get(), P.perm.get(), P.permSorted.get(), static_cast<unsigned>(n), bits, st);
  PermuteArgs A{};
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
  // |scale| = |q| |dx_a| / (dx dy dz dt) < |q| c dt / (dx dy dz dt), with a margin for rounding
  const double perCharge = kC * F.g.dt / (F.g.dx * F.g.dy * F.g.dz * F.g.dt) * (1.0 + 1e-9);
  const TileBoundsArgs T{P.keySorted.get(), n, ntiles, C.tileBegin.get(), C.tileEnd.get(), C.tileBound.get(),
                         perCharge};
  kPermute<<<dim3(blocks, 4), 256, 0, st>>>(A, P.permSorted.get(), n, T);
  launchCheck("permute");
  for (int a = 0; a < 9; ++a) P.s[a].swap(P.t[a]);
  P.id.swap(P.idT);
  P.state.swap(P.stateT);
  P.beta.swap(P.betaT);
  const int rz = C.shape == kShapeQuintic ? Box<kShapeQuintic>::RZ : Box<kShapeBinomial3X>::RZ;
  buildChunkLists(C, st, ntx, nty, ntiles, rz, C.mailbox.device(), seq);
}

// From the tile ranges of the sorted keys (written by the permute): packed
// chunk, occupied and active counts, their scan, the chunk list and the
// active-tile list (all in tile order).
void buildChunkLists(CouplingState& C, hipStream_t st, int ntx, int nty, int ntiles, int rz, unsigned* box,
                     unsigned seq) {
  const int ntz = ntiles / (ntx * nty);
  kTileCounts<<<ceilDiv(ntiles + 1, 256), 256, 0, st>>>(C.tileBegin.get(), C.tileEnd.get(), ntx, nty, ntz, rz,
                                                         C.tileCounts.get());
  launchCheck("tile counts");
  exclusiveScan64(C.tileCounts.get(), C.tileScan.get(), static_cast<std::size_t>(ntiles) + 1, C.scanTemp.get(),
                  C.scanTempBytes, st);
  kChunkList<<<ceilDiv(ntiles, 256), 256, 0, st>>>(C.tileBegin.get(), C.tileEnd.get(), C.tileCounts.get(),
                                                    C.tileScan.get(), ntiles, C.chunks.get(), C.activeTiles.get(),
                                                    box, seq, C.error.get());
  launchCheck("chunk list");
}

// Tile arrays sized for ntiles and a chunk list for n particles.
void reserveTileLists(CouplingState& C, hipStream_t st, int ntiles, int n) {
  if (static_cast<unsigned long long>(ntiles) > kMaxTiles)
    throw std::invalid_argument("The grid has more than " + std::to_string(kMaxTiles) + " cell tiles of 8x8x4");
  if (C.tileBegin.size() != static_cast<std::size_t>(ntiles)) {
    C.tileBegin.resize(ntiles);
    C.tileEnd.resize(ntiles);
    C.tileShift.resize(ntiles);
    C.activeTiles.resize(ntiles);
    C.tileBound.resize(ntiles);
    hipCheck(hipMemsetAsync(C.tileBound.get(), 0, static_cast<std::size_t>(ntiles) * sizeof(unsigned long long), st),
             "hipMemsetAsync");
    C.tileCounts.resize(ntiles + 1);
    C.tileScan.resize(ntiles + 1);
    C.scanTempBytes = std::max<std::size_t>(scanTempBytes64(static_cast<std::size_t>(ntiles) + 1), 256);
    C.scanTemp.resize(C.scanTempBytes);
    C.chunkCapacity = 0;
  }
  const std::size_t chunkBound = chunkListBound(ntiles, n);
  if (chunkBound > kMaxChunks)
    throw std::invalid_argument("More than " + std::to_string(kMaxChunks) + " particle chunks in one step");
  if (chunkBound > C.chunkCapacity) {
    // headroom: admissions raise the bound a little in many steps
    const std::size_t capacity = std::min<std::size_t>(chunkBound + chunkBound / 4 + 64, kMaxChunks);
    C.chunks.resize(capacity);
    C.chunkCapacity = capacity;
  }
}

// At most one partial chunk per occupied tile plus the full ones.
std::size_t chunkListBound(int ntiles, int n) {
  return (static_cast<std::size_t>(n) + kChunk - 1) / kChunk +
         std::min<std::size_t>(static_cast<std::size_t>(ntiles), static_cast<std::size_t>(n));
}

void fillInts(hipStream_t st, int* p, int n, int v) {
  kFillInt<<<ceilDiv(n ? n : 1, 256), 256, 0, st>>>(p, n, v);
  launchCheck("fill");
}

void reduceChunkStats(hipStream_t st, const double* in, int chunks, double* out) {
  kReduceBlocks<<<1, 256, 0, st>>>(in, chunks, out, nullptr, 0u, PicardWords{nullptr, 0u, nullptr, 0u});
  launchCheck("chunk stats");
}

void reserveStepBuffers(CouplingState& C, std::size_t partialNeed, unsigned nChunks, hipStream_t st) {
  if (partialNeed > C.partial.size()) {
    // tile partials must start at zero; the reduction keeps them so
    const std::size_t n = partialNeed + partialNeed / 2;
    C.partial.resize(n);
    hipCheck(hipMemsetAsync(C.partial.get(), 0, n * sizeof(long long), st), "hipMemsetAsync");
  }
  const std::size_t blocks = 4 * static_cast<std::size_t>(std::max(1u, nChunks));
  if (C.reduceBlocks.size() < blocks) C.reduceBlocks.resize(blocks + blocks / 2);
}

void reserveCouplingWords(CouplingState& C, hipStream_t st) {
  if (C.error.size() < static_cast<std::size_t>(kErrorWords)) C.error.resize(kErrorWords);
  if (C.reduceOut.size() < 4) C.reduceOut.resize(4);

  if (!C.counterQ.size()) {
    C.counterQ.resize(kCounterSlots);
    C.counterN.resize(kCounterSlots);
    hipCheck(hipMemsetAsync(C.counterQ.get(), 0, kCounterSlots * sizeof(double), st), "hipMemsetAsync");
    hipCheck(hipMemsetAsync(C.counterN.get(), 0, kCounterSlots * sizeof(unsigned long long), st), "hipMemsetAsync");
  }
}

namespace {
// Block partials of the counters, added in block order to the cumulative
// counters; a null flag always adds, otherwise only when it carries kEvtAny.
__global__ void __launch_bounds__(64) kCounterTotals(const double* blockQ, const unsigned* blockN, int blocks,
                                                     const unsigned* flag, double* q, unsigned long long* count) {
  if (flag && !(*flag & kEvtAny)) return;
  const int t = static_cast<int>(threadIdx.x);
  if (t >= kCounterSlots) return;
  double acc = 0.0;
  unsigned long long cnt = 0;
  for (int b = 0; b < blocks; ++b) {
    acc += blockQ[b * kCounterSlots + t];
    cnt += blockN[b * kCounterSlots + t];
  }
  q[t] += acc;
  count[t] += cnt;
}

void reserveBlockCounters(CouplingState& C, int blocks) {
  const std::size_t need = static_cast<std::size_t>(std::max(blocks, 1)) * kCounterSlots;
  if (C.blockQ.size() < need) {
    C.blockQ.resize(need + need / 2);
    C.blockN.resize(need + need / 2);
  }
}
}  // namespace

void commitEvents(CouplingState& C, hipStream_t st, int n, const int* evt, int* state, void* w, const void* q,
                  void* wend, unsigned* removed, const unsigned* trialWord, bool fp32) {
  if (n <= 0) return;
  const int blocks = static_cast<int>(ceilDiv(n, kEventBlock));
  reserveBlockCounters(C, blocks);
  if (fp32)
    kCommitEvents<float><<<blocks, kEventBlock, 0, st>>>(n, evt, state, static_cast<float*>(w),
                                                          static_cast<const float*>(q), static_cast<float*>(wend),
                                                          removed, trialWord, C.blockQ.get(), C.blockN.get());
  else
    kCommitEvents<double><<<blocks, kEventBlock, 0, st>>>(n, evt, state, static_cast<double*>(w),
                                                           static_cast<const double*>(q), static_cast<double*>(wend),
                                                           removed, trialWord, C.blockQ.get(), C.blockN.get());
  kCounterTotals<<<1, 64, 0, st>>>(C.blockQ.get(), C.blockN.get(), blocks, trialWord, C.counterQ.get(),
                                   C.counterN.get());
  launchCheck("commit events");
}

void cropCounters(CouplingState& C, hipStream_t st, int lo, int n, const int* state, const void* q, const void* w,
                  bool fp32) {
  if (n <= lo) return;
  const int blocks = static_cast<int>(ceilDiv(n - lo, kEventBlock));
  reserveBlockCounters(C, blocks);
  if (fp32)
    kCropCounters<float><<<blocks, kEventBlock, 0, st>>>(lo, n, state, static_cast<const float*>(q),
                                                          static_cast<const float*>(w), C.blockQ.get(), C.blockN.get());
  else
    kCropCounters<double><<<blocks, kEventBlock, 0, st>>>(lo, n, state, static_cast<const double*>(q),
                                                           static_cast<const double*>(w), C.blockQ.get(),
                                                           C.blockN.get());
  kCounterTotals<<<1, 64, 0, st>>>(C.blockQ.get(), C.blockN.get(), blocks, nullptr, C.counterQ.get(),
                                   C.counterN.get());
  launchCheck("crop counters");
}

std::size_t liveEntries(std::size_t count, const DeviceArray<unsigned>& removed, hipStream_t st) {
  if (!count || !removed.size()) return count;
  unsigned r = 0;
  hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
  hipCheck(hipMemcpy(&r, removed.get(), sizeof(unsigned), hipMemcpyDeviceToHost), "hipMemcpy");
  return count - std::min<std::size_t>(count, r);
}

// The step's sort phase on stream ss: tile lists and readback words
// sized, error words cleared, then (with particles) the stable tile sort,
// the state permute and the chunk lists, whose totals are posted to the
// mailbox under the returned sequence number.
static unsigned sortPhase(CouplingState& C, ParticleStore& P, const double origin[3], hipStream_t ss) {
  FieldState& F = *C.fields;
  const int n = static_cast<int>(P.count);
  const int ntx = static_cast<int>((F.g.nx + kTileX - 1) / kTileX);
  const int nty = static_cast<int>((F.g.ny + kTileY - 1) / kTileY);
  const int ntz = static_cast<int>((F.g.nz + kTileZ - 1) / kTileZ);
  const int ntiles = ntx * nty * ntz;
  reserveTileLists(C, ss, ntiles, n);
  reserveCouplingWords(C, ss);
  kFillInt<<<1, 128, 0, ss>>>(reinterpret_cast<int*>(C.error.get()), kErrorWords, 0);
  launchCheck("fill");
  if (!n) return 0;
  SortArgs K{};
  K.window = F.window ? 1 : 0;
  K.nz = static_cast<int>(F.g.nz);
  K.lowZ = origin[2];
  K.upZ = std::fma(static_cast<double>(F.g.nz), F.g.dz, origin[2]);
  K.dz = F.g.dz;
  K.words = C.error.get();
  K.removed = P.removed.get();
  const unsigned seq = C.mailbox.next(0);
  sortAndChunk(C, P, gatherArgs(C, F, origin), K, ntx, nty, ntiles, seq, ss);
  return seq;
}

bool preparedFor(const CouplingState& C, const void* store, std::uint64_t generation, std::size_t count,
                 const double origin[3]) {
  const auto& R = C.prepared;
  return R.valid && R.store == store && R.generation == generation && R.count == count && R.origin[0] == origin[0] &&
         R.origin[1] == origin[1] && R.origin[2] == origin[2];
}

// Starts the next step's sort phase on the coupling's sort stream after
// everything already enqueued on the grid's stream, so that it overlaps
// the field work enqueued after it (clear, window begin, magnetic update).
void prepareStep(CouplingState& C, ParticleStore& P, const double origin[3]) {
  FieldState& F = *C.fields;
  if (C.prepared.valid) throw std::logic_error("A prepared coupling step is already pending");
  hipStream_t ss = C.sortStream.get();
  // J of the last deposit is cleared through that step's active-tile list,
  // which this sort rewrites: clear it now, ahead of the sort (every step
  // starts from J = 0, and the step's own clear then finds nothing to do)
  if (F.footprint.valid && !F.currentZero) launchClearCurrent(F);
  hipCheck(hipEventRecord(C.prepareStart.get(), F.stream), "hipEventRecord");
  hipCheck(hipStreamWaitEvent(ss, C.prepareStart.get(), 0), "hipStreamWaitEvent");
  const unsigned seq = sortPhase(C, P, origin, ss);
  hipCheck(hipEventRecord(C.sortReady.get(), ss), "hipEventRecord");
  auto& R = C.prepared;
  R.valid = true;
  R.store = &P;
  R.generation = P.generation;
  R.count = P.count;
  for (int a = 0; a < 3; ++a) R.origin[a] = origin[a];
  R.seq = seq;
}

// Per-step state of a plain FP64 step between its two phases.
struct PlainStepContext {
  bool pending = false;
  int n = 0, ntx = 0, nty = 0, ntz = 0;
  unsigned nChunks = 0, nActive = 0, occupied = 0, activeEntries = 0;
  unsigned enqueued = 0;       // Picard trials enqueued so far
  unsigned checkedTrials = 0;  // trial words read back so far
  unsigned postSeq = 0;        // mailbox sequence of the last posted trial
  unsigned commitAfter = 0;    // trials ahead of the gated E commit (0: none enqueued)
  bool quintic = false;
  double origin[3] = {0.0, 0.0, 0.0};
  PArgs A{};
  GArgs G{};
  StepArgs S{};
};

namespace {

// Enqueue Picard trials [first, last) of the step; the last one posts its
// residual, the stop word and the unchecked trial words to the mailbox.
// Every trial runs behind the device stop word: once a trial has converged
// (or failed), the kernels of the later ones return at once, so trials can
// be enqueued ahead of the convergence decision.
void enqueueTrials(CouplingState& C, PlainStepContext& X, unsigned first, unsigned last) {
  FieldState& F = *C.fields;
  hipStream_t st = F.stream;
  const int n = X.n, ntx = X.ntx, nty = X.nty, ntz = X.ntz;
  const unsigned nChunks = X.nChunks, nActive = X.nActive;
  const bool quintic = X.quintic;
  PArgs& A = X.A;
  const GArgs& G = X.G;
  const StepArgs& S = X.S;
  unsigned* const stop = C.error.get() + kStopWord;
  for (unsigned trial = first; trial < last; ++trial) {
    unsigned* errT = C.error.get() + kTrialWord + trial;
    if (nChunks && trial == 0) {
      // The first trial pushes against the preview with J cleared
      // (simulation_source_coupling.cpp:43).
      GArgs G0 = G;
      for (int a = 0; a < 3; ++a) G0.j[a] = nullptr;
      G0.stop = nullptr;
      ScopedKernel sk(F.profiler, st, "gather");
      if (quintic)
        kGatherChunks<0, kShapeQuintic><<<nChunks, 2 * kChunk, 0, st>>>(A, G0, S, C.chunks.get(), ntx, nty, F.g.dt,
                                                                         nullptr, errT);
      else
        kGatherChunks<0, kShapeBinomial3X><<<nChunks, 2 * kChunk, 0, st>>>(A, G0, S, C.chunks.get(), ntx, nty, F.g.dt,
                                                                            nullptr, errT);
      launchCheck("gather");
    }
    if (n) {
      ScopedKernel sk(F.profiler, st, "push");
      kPush<<<ceilDiv(n, 128), 128, 0, st>>>(A, G, S, errT);
      launchCheck("push");
    }
    if (nChunks) {
      {
        ScopedKernel sk(F.profiler, st, "deposit");
        if (quintic)
          kDeposit<kShapeQuintic><<<nChunks, kChunk * depositLanes<kShapeQuintic>(), 0, st>>>(
              A, G, S, C.chunks.get(), C.tileBegin.get(), C.tileEnd.get(), C.tileScan.get(), ntx, nty,
              C.partial.get(), C.tileShift.get(), errT);
        else
          kDeposit<kShapeBinomial3X><<<nChunks, kChunk * depositLanes<kShapeBinomial3X>(), 0, st>>>(
              A, G, S, C.chunks.get(), C.tileBegin.get(), C.tileEnd.get(), C.tileScan.get(), ntx, nty,
              C.partial.get(), C.tileShift.get(), errT);
        launchCheck("deposit");
      }
      {
        ScopedKernel sk(F.profiler, st, "reduce_current");
        if (quintic)
          kReduceCurrent<kShapeQuintic><<<nActive, 256, 0, st>>>(F.base(6), F.base(7), F.base(8), G,
                                                                 C.activeTiles.get(), C.tileScan.get(),
                                                                 C.tileShift.get(), C.partial.get(), ntx, nty, ntz);
        else
          kReduceCurrent<kShapeBinomial3X><<<nActive, 256, 0, st>>>(F.base(6), F.base(7), F.base(8), G,
                                                                    C.activeTiles.get(), C.tileScan.get(),
                                                                    C.tileShift.get(), C.partial.get(), ntx, nty,
                                                                    ntz);
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
    }
    // TransverseReferenceCurrent::apply after the trial's deposit
    // (simulation_source_coupling.cpp:44)
    referenceCurrentTrial(C, F, C.esum[2].get(), C.estar[2].get(), true, nChunks > 0, ntx, nty, ntz, stop);
    const bool post = trial + 1 == last;
    if (post) X.postSeq = C.mailbox.next(1);
    const PicardWords W{stop, trial, C.error.get() + kTrialWord, X.checkedTrials};
    unsigned* const box = post ? C.mailbox.device() : nullptr;
    const unsigned seq = post ? X.postSeq : 0u;
    if (nChunks) {
      ScopedKernel sk(F.profiler, st, "residual");
      if (quintic)
        kGatherChunks<1, kShapeQuintic><<<nChunks, 2 * kChunk, 0, st>>>(A, G, S, C.chunks.get(), ntx, nty, F.g.dt,
                                                                         C.reduceBlocks.get(), errT);
      else
        kGatherChunks<1, kShapeBinomial3X><<<nChunks, 2 * kChunk, 0, st>>>(A, G, S, C.chunks.get(), ntx, nty, F.g.dt,
                                                                            C.reduceBlocks.get(), errT);
    }
    // a separate launch: folding it into the gather's last workgroup needs
    // a device-scope fence in every workgroup (the XCDs' L2s write back),
    // which was measured to make the gather three times slower
    kReduceBlocks<<<1, 256, 0, st>>>(C.reduceBlocks.get(), static_cast<int>(nChunks), C.reduceOut.get(), box, seq, W);
    launchCheck("residual");
    // The next trial pushes with the fields this residual gathered.
    for (int a = 0; a < 6; ++a) std::swap(A.fin[a], A.fout[a]);
  }
  X.enqueued = last;
}

}  // namespace

void plainStepBegin(CouplingState& C, ParticleStore& P, const double origin[3], double time, unsigned substeps) {
  FieldState& F = *C.fields;
  if (!C.plain) C.plain = std::make_shared<PlainStepContext>();
  PlainStepContext& X = *C.plain;
  if (X.pending) throw std::logic_error("The previous coupling step was not finished");
  if (F.window && F.phase != 2) throw std::logic_error("Moving-window electric phase mismatch");
  if (F.window && !F.inletSet) throw std::logic_error("Moving-window inlet planes were not supplied");
  if (substeps == 0) throw std::invalid_argument("Particle substeps must be positive");
  int n = static_cast<int>(P.count);
  hipStream_t st = F.stream;
  // J is zero at the start of every step (simulation.cpp:518); the tile
  // reduction below rewrites only the neighbourhood of occupied tiles.
  launchClearCurrent(F);
  for (int a = 0; a < 3; ++a)
    if (C.estar[a].size() != F.compSize(a)) {
      C.estar[a].resize(F.compSize(a));
      C.esum[a].resize(F.compSize(a));
    }

  const int ntx = static_cast<int>((F.g.nx + kTileX - 1) / kTileX);
  const int nty = static_cast<int>((F.g.ny + kTileY - 1) / kTileY);
  const int ntz = static_cast<int>((F.g.nz + kTileZ - 1) / kTileZ);
  const int ntiles = ntx * nty * ntz;
  const std::size_t chunkBound = chunkListBound(ntiles, n);
  // Readbacks arrive in the mailbox: the scan total (chunks, active tiles)
  // and the sort-phase words at words 8..14 (flag 0); the posting trial's
  // residual sums (doubles), the stop word and the trial error words from
  // word 15 (flag 1).
  unsigned sortSeq = 0;
  if (preparedFor(C, &P, P.generation, P.count, origin)) {
    hipCheck(hipStreamWaitEvent(st, C.sortReady.get(), 0), "hipStreamWaitEvent");
    sortSeq = C.prepared.seq;
  } else {
    if (C.prepared.valid) hipCheck(hipStreamWaitEvent(st, C.sortReady.get(), 0), "hipStreamWaitEvent");
    sortSeq = sortPhase(C, P, origin, st);
  }
  C.prepared.valid = false;
  // Zero-current preview E* (MovingWindowFields::previewElectricBoundary);
  // it runs while the host waits for the chunk count.
  // E* and E* + E^n are read only on the nodes of active tiles (gathers
  // stage the boxes of occupied tiles; the reduction rewrites active tiles).
  double* const out[3] = {C.estar[0].get(), C.estar[1].get(), C.estar[2].get()};
  double* const sums[3] = {C.esum[0].get(), C.esum[1].get(), C.esum[2].get()};
  if (n)
    launchElectricPreviewTiles(F, F.ptrs(), out, sums, F.window ? F.inletPlane[0] : nullptr,
                               F.window ? F.inletPlane[1] : nullptr, C.activeTiles.get(), C.tileScan.get() + ntiles,
                               TileGrid{ntx, nty, ntz, kTileX, kTileY, kTileZ});
  unsigned nChunks = 0, nActive = 0, occupiedTiles = 0, activeEntries = 0;
  if (n) {
    C.mailbox.wait(0, sortSeq, st);
    const unsigned* box = C.mailbox.host();
    nChunks = box[8];
    nActive = box[9];
    occupiedTiles = box[11];
    // finishMovingWindowStep's checks of the entries, applied to this origin
    raiseParticleErrors(box[10]);
    activeEntries = box[14];
    const int behind = static_cast<int>(box[12]);
    if (box[13]) cropCounters(C, st, n - behind, n, P.state.get(), P.s[6].get(), P.s[8].get(), false);
    n -= behind;  // removed and cropped entries now trail the live ones
    P.count = static_cast<std::size_t>(n);
    if (nChunks > chunkBound) throw std::logic_error("Chunk list exceeds its bound");
  }
  const bool quintic = C.shape == kShapeQuintic;
  const int boxN = quintic ? boxNodes<kShapeQuintic>() : boxNodes<kShapeBinomial3X>();
  const std::size_t partialNeed = static_cast<std::size_t>(occupiedTiles) * 3 * boxN;
  reserveStepBuffers(C, partialNeed, nChunks, st);

  X.S = stepArgsFor(C, F.g, F.hasCpml ? &F.cpml : nullptr, F.window, origin, time, substeps);
  X.A = particleArgs(P);
  X.A.scaleDen = F.g.dx * F.g.dy * F.g.dz * F.g.dt;
  X.A.tileOf = P.keySorted.get();
  X.A.tileBound = C.tileBound.get();
  X.G = gatherArgs(C, F, origin);
  X.G.stop = C.error.get() + kStopWord;
  X.n = n;
  X.ntx = ntx;
  X.nty = nty;
  X.ntz = ntz;
  X.nChunks = nChunks;
  X.nActive = nActive;
  X.occupied = occupiedTiles;
  X.activeEntries = activeEntries;
  X.quintic = quintic;
  X.enqueued = 0;
  X.checkedTrials = 0;
  for (int a = 0; a < 3; ++a) X.origin[a] = origin[a];
  X.pending = true;
  // With active particles the convergence test needs at least two trials
  // (simulation_source_coupling.cpp:50); the first batch is the last
  // step's trial count.
  const unsigned batch = C.picardFixed ? C.picardFixed
                                       : std::min<unsigned>(kMaxTrials, std::max(C.picardBatch, activeEntries ? 2u : 1u));
  enqueueTrials(C, X, 0, batch);
  // commitConservativeElectric enqueued behind the batch, gated on the stop
  // word: it runs when a trial of the batch has converged, so the device
  // does not wait for the host to read the decision. (The energy ledger's
  // commit, several launches, stays after the decision, and so does the
  // commit while its workgroup size is being tuned: a skipped launch would
  // be a false sample.)
  X.commitAfter = 0;
  if (!(F.window && F.ledger.enabled) && autotuneSettled(shapeKey("electric", F.g, sizeof(double)))) {
    double* const none[3] = {nullptr, nullptr, nullptr};
    double* const* noSum = nullptr;
    launchElectric(F, F.ptrs(), 0, none, F.inletPlane[0], F.inletPlane[1], noSum, C.error.get() + kStopWord);
    X.commitAfter = X.enqueued;
  }
}

void plainStepFinish(CouplingState& C, ParticleStore& P, aprl_step_report* report) {
  if (!C.plain || !C.plain->pending) throw std::logic_error("Coupling step phase mismatch");
  PlainStepContext& X = *C.plain;
  X.pending = false;
  FieldState& F = *C.fields;
  hipStream_t st = F.stream;
  unsigned iterations = 0, lastWord = 0, stop = 0;
  double force = 0., work = 0.;
  for (;;) {
    C.mailbox.wait(1, X.postSeq, st);
    const unsigned* box = C.mailbox.host();
    const double* sums = C.mailbox.hostDoubles();
    stop = box[15];
    const unsigned ran = stop ? (stop & ~kStopError) : X.enqueued;
    // device errors in trial order, as the reference raises them
    for (unsigned t = X.checkedTrials; t < ran; ++t) raiseParticleErrors(box[16 + t - X.checkedTrials]);
    if (ran) lastWord = box[16 + ran - 1 - X.checkedTrials];
    X.checkedTrials = X.enqueued;
    iterations = ran;
    force = sums[0];
    work = sums[3] == 0. ? 0. : (sums[2] > 0. ? sums[1] / sums[2] : sums[1]);
    if (stop || X.enqueued == static_cast<unsigned>(kMaxTrials)) break;
    // not converged within the batch: one more trial at a time
    enqueueTrials(C, X, X.enqueued, X.enqueued + 1);
  }
  if (!stop) {
    std::string m = "Conservative source coupling failed: force=" + std::to_string(force) +
                    " work=" + std::to_string(work) + " iterations=" + std::to_string(iterations);
    throw std::runtime_error(m);
  }
  C.picardBatch = iterations;
  // commitConservativeElectric: E^n + curl B - dt/eps0 J, then the committed
  // CPML-E (history advance), PEC and inlet overwrite; done already by the
  // gated commit when the converged trial was ahead of it
  const bool gatedRan = X.commitAfter && iterations <= X.commitAfter && !(stop & kStopError);
  if (!gatedRan) {
    double* const none[3] = {nullptr, nullptr, nullptr};
    if (F.window) windowCommitElectric(F, F.ptrs());
    else launchElectric(F, F.ptrs(), 0, none, F.inletPlane[0], F.inletPlane[1]);
  }
  if (F.window) F.phase = 3;
  // Boundary events of the converged trial: carriers, removals, counters.
  if (lastWord & kEvtAny)
    commitEvents(C, st, X.n, P.evt.get(), P.state.get(), P.s[8].get(), P.s[6].get(), P.wend.get(), P.removed.get(),
                 C.error.get() + kTrialWord + (iterations - 1), false);
  // The trial end state becomes the particle state; the start positions
  // stay in the end arrays for the step diagnostics.
  for (int a = 0; a < 6; ++a) P.s[a].swap(P.end[a]);
  C.lastStep.valid = true;
  C.lastStep.store = &P;
  C.lastStep.generation = P.generation;
  for (int a = 0; a < 3; ++a) C.lastStep.origin[a] = X.origin[a];
  if (report) {
    report->iterations = iterations;
    report->force = force;
    report->work = work;
    report->tiles = X.occupied;
    report->chunks = X.nChunks;
    report->active_tiles = X.nActive;
  }
}

void couplingStep(CouplingState& C, ParticleStore& P, const double origin[3], double time, unsigned substeps,
                  aprl_step_report* report) {
  plainStepBegin(C, P, origin, time, substeps);
  plainStepFinish(C, P, report);
}

}  // namespace aprl_gpu
