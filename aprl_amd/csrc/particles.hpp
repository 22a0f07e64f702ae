// Device particle store and the conservative-source coupling state.
#pragma once

#include <memory>

#include "fields.hpp"
#include "tilecounts.hpp"

namespace aprl_gpu {

// Cell tile used for the per-step sort, the gathers and the deposit.
// Chords move less than one cell per axis per step (c*dt = dz <= dx, dy),
// so each shape's taps of a chord that starts in a tile lie in a fixed
// extended node box around the tile.
constexpr int kTileX = 8, kTileY = 8, kTileZ = 4;

// Error bits raised by kernels; the host maps the lowest set bit onto the
// reference's exception type and message.
enum ErrorBit : unsigned {
  kErrEndpointOutside = 1u << 0,   // eb_deposition.cpp:244-245 (out_of_range)
  kErrCurrentY = 1u << 1,          // longitudinal_shape.h:185 (out_of_range)
  kErrCurrentX = 1u << 2,          // longitudinal_shape.h:181-184 (out_of_range)
  kErrOrbitOutside = 1u << 3,      // source_coupling.cpp:78 (out_of_range)
  kErrMagneticOutside = 1u << 4,   // source_coupling.cpp:87 (out_of_range)
  kErrRootNonphysical = 1u << 5,   // source_coupling_math.h:285 (runtime_error)
  kErrRootNoConverge = 1u << 6,    // source_coupling_math.h:289 (runtime_error)
  kErrNonfinite = 1u << 7,         // source_coupling_math.h:241 (runtime_error)
  kErrRemapOutside = 1u << 8,      // simulation_moving_window.cpp:42, 358-360 (runtime_error)
  kErrStartInCpml = 1u << 9,       // particle_boundary.cpp:118-121 (out_of_range)
  kErrStartOutside = 1u << 10,     // particle_boundary.cpp:325-330 (out_of_range)
  kErrBeyondExit = 1u << 11,       // simulation_moving_window.cpp:387 (runtime_error)
  kErrUpstream = 1u << 12,         // simulation.cpp:4926-4930 (runtime_error)
  kErrCrossings = 1u << 13,        // orbit crossing budget (runtime_error)
  kErrTileHalo = 1u << 14,         // chord longer than one cell per axis (runtime_error)
  kErrMagneticNoDomain = 1u << 15, // source_coupling.cpp:84 (out_of_range)
  kErrCarrierBehind = 1u << 16,    // simulation_moving_window.cpp:365 (runtime_error)
  kErrCpmlLocate = 1u << 17,       // particle_boundary.cpp:153-157 (runtime_error)
  kErrExitLocate = 1u << 18,       // particle_boundary.cpp:356-360 (runtime_error)
  kErrBehindInlet = 1u << 19,      // simulation_moving_window.cpp:359 (runtime_error)
  kErrHeldRelease = 1u << 20,      // simulation.cpp:4966-4970 (runtime_error)
  kErrHeldFrozenLeg = 1u << 21,    // simulation.cpp:4687-4691 (runtime_error)
  kErrAdmission = 1u << 22,        // simulation.cpp:5141-5143 (runtime_error)
  // not an error: some entry ended a trial's segment at a boundary event
  kEvtAny = 1u << 31,
};

// State of a particle-store entry. A CPML carrier (ParticleCPMLCarrier,
// particle_boundary.h:76-84) keeps its proper velocity (velocity = u c /
// gamma) and holds its current weight in the weight slot. Removed and
// cropped entries stay in the arrays until the next step's sort moves them
// behind the live ones.
// A held entry (RelativisticParticleSI::frozen) moves z += c dt beta_held
// with its transverse motion held until its lab z reaches the activation
// plane; beta_held is kept per entry.
enum EntryState : int { kStateActive = 0, kStateCarrier = 1, kStateHeld = 2, kStateRemoved = -1, kStateCropped = -2 };

// Boundary counter slots (cumulative per coupling): CPML entries per face
// (0-5, faces x- x+ y- y+ z- z+), outer exits of particles (6-11) and of
// carriers (12-17), moving-window exits (18) and cropped carriers (19);
// simulation.h cpmlEntryCount_ / directOuterCount_ / carrierOuterCount_ and
// the window state's exits / croppedCarriers, with their charges.
constexpr int kCounterSlots = 20;

// Error words of a coupling step: [0] sort-phase errors, [1] unused (the
// occupied tiles come with the tile scan's total), [2] entries keyed
// behind the live ones (removed or cropped),
// [3] cropped carriers, [4] live active (not held) particles at the sort,
// [kTrialWord + t] the kernels of Picard trial t, [kStopWord] the plain
// form's Picard state: 0 while trials run, else the iteration count of the
// trial that converged or failed (kStopError); the trials' kernels return
// at once when it is set.
constexpr int kTrialWord = 5;
constexpr int kMaxTrials = 64;  // simulation_source_coupling.cpp:40
constexpr int kStopWord = kTrialWord + kMaxTrials;
constexpr int kErrorWords = kStopWord + 1;
constexpr unsigned kStopError = 1u << 31;

struct MagnetDev {
  int type;  // 0 undulator, 1 dipole, 2 quadrupole
  double center[3];
  double length, period, peak, angle;
  int gaussianFringe;
  double cutoff, gradient, endRampPeriods, phase;
  double waveNumber, fringeExtent, cosA, sinA;
};
constexpr int kMaxMagnets = 8;

struct BoostDev {
  double gamma, beta, gammaBeta, c0, labT0, labZ0, boxZ0;
};

struct ParticleStore {
  std::size_t capacity = 0, count = 0;
  std::uint64_t generation = 0;  // bumped by every host upload
  // SoA state: x y z ux uy uz q m w, id, entry state, held speed
  DeviceArray<double> s[9], t[9];
  DeviceArray<uint64_t> id, idT;
  DeviceArray<int> state, stateT;
  DeviceArray<double> beta, betaT;
  // Per-step slots: fields[0] and fields[1] alternate between the fields a
  // trial pushes with (Ex Ey Ez Bx By Bz) and the residual gather's next
  // fields; end[6] = trial end state x1, u1; gridAbs = per-substep work sum.
  DeviceArray<double> fields[2][6], end[6], gridAbs;
  // per-step boundary events (rewritten by every trial's push): event
  // code, chord fractions t1 (end of the unit-weight part) and t2 (end of
  // the carrier part), the carrier part's deposit weight, and the carrier
  // end weight (after the commit: the entry's weight at the step start)
  DeviceArray<int> evt;
  DeviceArray<double> cut[3], wend;
  // entries removed by the last commit (read lazily between steps)
  DeviceArray<unsigned> removed;
  // sort
  DeviceArray<unsigned> key, keySorted, perm, permSorted;
  DeviceArray<unsigned char> sortTemp;
  std::size_t sortTempBytes = 0;
};

// FP32-split particle store: per axis the absolute lattice cell (int32,
// floor(x / h) with the lattice anchored at x = 0) and the FP32 offset in
// [0, 1) inside it, so that neither long runs nor the moving window lose
// position bits; momenta, charge, mass and weight in FP32.
struct SplitStore {
  std::size_t capacity = 0, count = 0;
  std::uint64_t generation = 0;  // bumped by every host upload
  DeviceArray<int> cell[3], cellT[3];
  DeviceArray<float> s[9], t[9];  // offset x y z, u x y z, q m w
  DeviceArray<uint64_t> id, idT;
  DeviceArray<int> state, stateT;
  DeviceArray<float> beta, betaT;  // held speed of held entries
  // per-step slots as in ParticleStore: alternating gathered fields, the
  // trial end state (cell1, offset1, u1) and the per-step work sum
  DeviceArray<float> fields[2][6], end[6], gridAbs;
  DeviceArray<int> endCell[3];
  DeviceArray<float> chord[3];  // trial displacement x1 - x0 in cells
  DeviceArray<int> evt;            // per-step boundary events, as in ParticleStore
  DeviceArray<float> cut[3], wend;
  DeviceArray<unsigned> removed;
  DeviceArray<unsigned> key, keySorted, perm, permSorted;
  DeviceArray<unsigned char> sortTemp;
  std::size_t sortTempBytes = 0;
};

struct SplitStepContext;  // per-step state between the phases of a split step
struct PlainStepContext;  // the same for a plain FP64 step

// Particles per work chunk: one chunk is a contiguous run of tile-sorted
// particles of one tile. The deposit gives each chunk one workgroup of
// kChunk lanes; the gathers give it 2*kChunk lanes (E waves and B waves).
constexpr int kChunk = 128;

struct CouplingState {
  FieldState* fields = nullptr;      // plain FP64 grid, or
  FieldState32* fields32 = nullptr;  // FP32-split grid
  int shape = 2;
  bool smoothX = true;
  BoostDev frame{1.0, 0.0, 0.0, kC, 0.0, 0.0, 0.0};
  std::vector<MagnetDev> magnets;
  std::size_t cpmlCells[3] = {0, 0, 0};
  double exitLabZ = 0.0;
  bool hasExit = false;
  // lab activation plane of held particles (config reference.activation_plane_z)
  double activationLabZ = 0.0;
  bool hasActivation = false;
  bool holdTransverse = true;  // reference.hold_transverse_until_activation
  DeviceArray<double> estar[3];
  // E_cand + E_old of the current trial (the preview writes the zero-current
  // sums on the whole grid; the current reduction rewrites the active tiles)
  DeviceArray<double> esum[3];
  DeviceArray<float> estar32[3], esum32[3];  // the same for an FP32-split grid
  // FP32-split form: Picard trials per step (no host round trip inside a
  // step) and whether a last residual gather reports force and work
  unsigned splitTrials = 2;
  bool splitVerify = false;
  // tiles: particle range, fixed-point exponent of the tile's current,
  // active-tile list
  DeviceArray<int> tileBegin, tileEnd, tileShift, activeTiles;
  // per tile: the largest current bound of one of its particles in the
  // trial (bit pattern of a non-negative double; written by the push,
  // read by the deposit, cleared by the current reduction and the sort)
  DeviceArray<unsigned long long> tileBound;
  // per tile (+1 entry for the scan totals): packed counts (chunks,
  // occupied, has an occupied neighbour; packTileCounts); the exclusive
  // scan gives the first chunk, the partial slot and the active slot
  DeviceArray<unsigned long long> tileCounts, tileScan;
  DeviceArray<int4> chunks;  // (tile, first particle, end particle, 0)
  std::size_t chunkCapacity = 0;
  DeviceArray<unsigned char> scanTemp;
  std::size_t scanTempBytes = 0;
  // per occupied tile (in tile order): 3 extended tile boxes of
  // fixed-point current
  DeviceArray<long long> partial;
  DeviceArray<double> reduceBlocks, reduceOut;
  DeviceArray<unsigned> error;
  // cumulative boundary counters and their per-block partials
  DeviceArray<double> counterQ, blockQ;
  DeviceArray<unsigned long long> counterN;
  DeviceArray<unsigned> blockN;
  PinnedHost pinned;
  HostMailbox mailbox;  // kernel-written readbacks of the plain form
  Event sorted, stepDone;
  std::shared_ptr<SplitStepContext> split;
  std::shared_ptr<PlainStepContext> plain;
  // plain form: Picard trials enqueued before the host first waits (the
  // last step's count; later trials of a converged step return at once)
  unsigned picardBatch = 2;
  unsigned picardFixed = 0;  // a fixed first batch (0: adaptive)
  // A sort phase started ahead of the step (prepareStep) on its own
  // stream, so that it overlaps the field updates enqueued after it:
  // sortStream waits for `prepareStart` (recorded on the grid's stream at
  // the prepare) and `sortReady` marks its end. The step uses it only for
  // the same store, generation, entry count and origin.
  Stream sortStream;
  Event prepareStart, sortReady;
  struct {
    bool valid = false;
    const void* store = nullptr;
    std::uint64_t generation = 0;
    std::size_t count = 0;
    double origin[3] = {0.0, 0.0, 0.0};
    unsigned seq = 0;
  } prepared;
  // reference surface current (TransverseReferenceCurrent): rows x points
  // table, target (i, j) pairs, translation beta
  bool refCurrent = false;
  DeviceArray<double> refTable;
  DeviceArray<int> refPoints;
  std::size_t refRows = 0, refCount = 0;
  double refBeta = 0.0;
  // the last coupled step: its store and origin (its start positions stay
  // in the store's end arrays until the next step)
  struct {
    bool valid = false;
    const void* store = nullptr;  // ParticleStore or SplitStore
    std::uint64_t generation = 0;
    double origin[3] = {0.0, 0.0, 0.0};
  } lastStep;
  DeviceArray<double> rho[2];  // diagnostic charge densities
};

// Raise a tile's current bound to b (b >= 0): the lanes of a wave that
// hold the same tile (sorted particles: usually the whole wave) combine
// first, so that one atomic per tile and wave reaches memory. Every lane of
// the wave must call it; lanes without a particle pass tile = ~0u.
__device__ __forceinline__ void raiseTileBound(unsigned long long* tileBound, unsigned tile, double b) {
  const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(b));
  bool pending = tile != ~0u;
  const int lane = static_cast<int>(threadIdx.x & 63u);
  while (true) {
    const unsigned long long mask = __ballot(pending);
    if (!mask) break;
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    const unsigned lt = __shfl(tile, leader, 64);
    unsigned long long v = pending && tile == lt ? bits : 0ull;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(v, off, 64);
      v = o > v ? o : v;
    }
    if (lane == leader) atomicMax(tileBound + lt, v);
    if (tile == lt) pending = false;
  }
}

// Fixed-point exponent of a tile: every tap of the tile's particles is at
// most weightBound x (largest per-particle bound), so the tile's integer
// sums stay below 2^62 with count particles.
__device__ __forceinline__ int tileFixedShift(double weightBound, unsigned long long boundBits, int count) {
  const double b = __longlong_as_double(static_cast<long long>(boundBits));
  int e2 = 0;
  (void)frexp(weightBound * b * static_cast<double>(count), &e2);
  return 62 - e2;
}

// Run boundaries of the sorted tile keys, folded into the permute kernels:
// entry p opens its tile's range when it differs from its predecessor and
// closes it when it differs from its successor.
struct TileBoundsArgs {
  const unsigned* key;  // sorted keys
  int n, ntiles;
  int *begin, *end;
  unsigned long long* bound;  // per tile: largest entry current bound (cleared by the key kernels)
  double perCharge;           // current bound per unit |charge| of a step's chord
};

// An entry's current bound for the step, known at the sort: every substep
// moves a particle or a carrier less than c dt along any axis (|v| < c), so
// a chord's current scale stays below |q| c dt / (dx dy dz dt) (perCharge,
// with a rounding margin); a held entry deposits at most its held leg and
// the released remainder, twice that.
__device__ __forceinline__ double entryCurrentBound(double q, int state, double perCharge) {
  return fabs(q) * perCharge * (state == kStateHeld ? 2.0 : 1.0);
}

__device__ __forceinline__ void tileBoundsAt(const TileBoundsArgs& T, int p) {
  const unsigned k = T.key[p];
  if (k >= static_cast<unsigned>(T.ntiles)) return;  // removed or cropped entries sort last
  if (p == 0 || T.key[p - 1] != k) T.begin[k] = p;
  if (p == T.n - 1 || T.key[p + 1] != k) T.end[k] = p + 1;
}

// Tile ranges cleared to empty by the key kernels (grid-stride over the
// tiles), so that the boundary pass after the sort only writes occupied
// tiles.
struct TileRangeArgs {
  int* begin;
  int* end;
  int ntiles;
  unsigned long long* bound;
};
__device__ __forceinline__ void clearTileRanges(const TileRangeArgs& R, int gid) {
  for (int t = gid; t < R.ntiles; t += static_cast<int>(gridDim.x * blockDim.x)) {
    R.begin[t] = 0;
    R.end[t] = 0;
    R.bound[t] = 0ull;
  }
}

// Host helpers shared by the FP64 and FP32-split coupling steps.
// box / seq: the mailbox (device address) that receives the scan total and
// the sort-phase words, or nullptr.
void buildChunkLists(CouplingState& C, hipStream_t st, int ntx, int nty, int ntiles, int rz, unsigned* box = nullptr,
                     unsigned seq = 0);
void reserveTileLists(CouplingState& C, hipStream_t st, int ntiles, int n);
std::size_t chunkListBound(int ntiles, int n);
void fillInts(hipStream_t st, int* p, int n, int v);
void reduceChunkStats(hipStream_t st, const double* in, int chunks, double* out);
void raiseParticleErrors(unsigned e);
void reserveParticles(ParticleStore& P, std::size_t capacity);
// Appends k host entries after the store's count (growing it and keeping
// what it holds): positions and momenta as xyz triples, per-entry entry
// state and held speed.
void appendParticles(ParticleStore& P, std::size_t k, const double* x3, const double* u3, const double* q,
                     const double* m, const double* w, const uint64_t* id, const int* state, const double* beta);
// Error words and boundary counters of a coupling (allocated and cleared
// once).
void reserveCouplingWords(CouplingState& C, hipStream_t st);
// The step's tile partials (zeroed) and residual block slots, grown with
// half again as much headroom: a bunch that spreads over new tiles raises
// the need a little in many steps, and every reallocation frees device
// memory, which waits for the device.
void reserveStepBuffers(CouplingState& C, std::size_t partialNeed, unsigned nChunks, hipStream_t st);
// Boundary-event commit after the converged trial: entry states and
// weights, removed-entry count and the counters, skipped on the device when
// the trial's error word carries no kEvtAny. R is the store's real type.
void commitEvents(CouplingState& C, hipStream_t st, int n, const int* evt, int* state, void* w, const void* q,
                  void* wend, unsigned* removed, const unsigned* trialWord, bool fp32);
// Cropped-carrier counters of the sorted tail [lo, n).
void cropCounters(CouplingState& C, hipStream_t st, int lo, int n, const int* state, const void* q, const void* w,
                  bool fp32);
// Reference surface current of one trial on the Jz ring nodes: Jz = (the
// trial's particle current on active tiles, else 0) + the interpolated
// table value; with writeEsum the E_cand + E_old sums of the active ring
// nodes follow. Checks the table range and leaves J for a full clear.
// A non-null stop word skips the update once it is set (the plain form's
// speculative trials).
template <class R>
void referenceCurrentTrial(CouplingState& C, FieldStateT<R>& F, R* esumZ, const R* estarZ, bool writeEsum,
                           bool anyActive, int ntx, int nty, int ntz, const unsigned* stop = nullptr);
// Live entries of a store whose last commit may have removed some.
std::size_t liveEntries(std::size_t count, const DeviceArray<unsigned>& removed, hipStream_t st);
// Stable sort of (tile key, particle index) pairs on the low `bits` key
// bits; temp == nullptr returns the scratch size in bytes.
void tileSort(void* temp, std::size_t& bytes, unsigned* key, unsigned* keySorted, unsigned* perm,
              unsigned* permSorted, unsigned n, int bits, hipStream_t st);
// Device-wide exclusive scans (out may equal in); the scratch sizes are in
// bytes (0: no scratch needed).
std::size_t scanTempBytes32(std::size_t n);
std::size_t scanTempBytes64(std::size_t n);
void exclusiveScan32(const unsigned* in, unsigned* out, std::size_t n, void* temp, std::size_t bytes,
                     hipStream_t st);
void exclusiveScan64(const unsigned long long* in, unsigned long long* out, std::size_t n, void* temp,
                     std::size_t bytes, hipStream_t st);

// FP32-split form (particles32.hip).
void reserveSplitParticles(SplitStore& P, std::size_t capacity);
void appendSplitParticles(SplitStore& P, const FieldState32& F, std::size_t k, const double* x3, const double* u3,
                          const double* q, const double* m, const double* w, const uint64_t* id, const int* state,
                          const double* beta);
void uploadSplitParticles(SplitStore& P, const FieldState32& F, std::size_t n, const double* x3, const double* u3,
                          const double* q, const double* m, const double* w, const uint64_t* id);
void downloadSplitParticles(const SplitStore& P, const FieldState32& F, double* x3, double* u3, double* q, double* m,
                            double* w, uint64_t* id);
void couplingStepSplit(CouplingState& C, SplitStore& P, const double origin[3], double time, unsigned substeps,
                       aprl_step_report* report);
// The sort phase of the next coupled step started ahead of it (before the
// field updates of the step are enqueued) on the coupling's sort stream;
// the step uses it when called with the same store, unchanged since, and
// the same origin, and otherwise sorts again.
void prepareStep(CouplingState& C, ParticleStore& P, const double origin[3]);
void prepareStepSplit(CouplingState& C, SplitStore& P, const double origin[3]);
bool preparedFor(const CouplingState& C, const void* store, std::uint64_t generation, std::size_t count,
                 const double origin[3]);
// The same in three phases, so that steps of independent instances can be
// interleaved: begin (sort, preview), middle (trials, commit, async
// readback), end (wait, errors, report).
void splitStepBegin(CouplingState& C, SplitStore& P, const double origin[3], double time, unsigned substeps);
// Plain FP64 step in two phases: begin (sort or the prepared sort, preview,
// the first batch of Picard trials) and finish (wait for the device's
// convergence decision, further trials if needed, commit, events, report).
void plainStepBegin(CouplingState& C, ParticleStore& P, const double origin[3], double time, unsigned substeps);
void plainStepFinish(CouplingState& C, ParticleStore& P, aprl_step_report* report);
void splitStepMiddle(CouplingState& C, SplitStore& P);
void splitStepEnd(CouplingState& C, SplitStore& P, aprl_step_report* report);

}  // namespace aprl_gpu
