// The laboratory-plane input route of a moving-window run on the device:
// plane-arrival events of a seeded Gaussian bunch and the reference's
// startup sequence for them (Simulation's initialization,
// simulation.cpp:1983-2245) as one lane per macro-particle:
//   SIBunchPreprocessor::projectLabPlaneEventsToSnapshot (eb_bunch.cpp:134-289)
//   SIBunchPreprocessor::placeRelativeLabSnapshot        (eb_bunch.cpp:291-352)
//   the frozen-injection frame origin                    (simulation.cpp:2126-2186)
//   SIBunchPreprocessor::boostLabSnapshotToBoxTimeZero   (eb_bunch.cpp:354-482)
//   the frozen flags and held speeds                     (simulation.cpp:2228-2245)
// with BoostFrameTransform's light-front momentum transforms
// (boostframe.cpp:74-164). The per-particle arithmetic follows the
// reference's order. The three sums the reference keeps in long double are
// double-double sums here (fixed block order), finished on the host in long
// double; minima and maxima are exact (order-preserving atomics).
#include <algorithm>
#include <cmath>
#include <iomanip>
#include <limits>
#include <memory>
#include <sstream>
#include <vector>

#include "ddsum.hpp"
#include "gaussian_draw.hpp"
#include "handles.hpp"

namespace aprl_gpu {
namespace {

using namespace ddsum;

constexpr int kThreads = 256, kMaxBlocks = 1024;
constexpr double kEps = std::numeric_limits<double>::epsilon();

// error bits raised by the kernels (host messages in raiseRouteErrors)
enum : unsigned {
  kErrEventInput = 1u << 0,      // projection: weight, arrival time or beta_z invalid
  kErrPositionFinite = 1u << 1,  // placement: non-finite z
  kErrHeldAdvance = 1u << 2,     // frame origin: held pseudo-trajectory does not advance
  kErrHeldVelocity = 1u << 3,    // boost: held mode needs positive lab beta_z
  kErrHeldLab = 1u << 4,         // boost: held pseudo-trajectory does not advance in lab z
  kErrSimultaneity = 1u << 5,    // boost: no intersection with the box-time-zero plane
  kErrBoxTimeZero = 1u << 6,     // boost: free drift missed box time zero
  kErrTransverseMass = 1u << 7,  // light front: transverse mass not finite
  kErrLightFront = 1u << 8,      // light front: invalid positive component
  kErrLongitudinal = 1u << 9,    // light front: non-finite u_z
  kErrAlreadyPast = 1u << 10,    // freeze: a particle at or past the activation plane
};

struct Bunch {
  double* x[3];
  double* u[3];
  double* w;
  double* beta;
  long long n;
};

// BoostFrameTransform with c0 = c: gamma, beta, gamma beta and the origins.
struct FrameDev {
  double gamma, beta, gammaBeta, labT0, labZ0, boxZ0;
};

FrameDev frameFrom(double gamma, double labT0, double labZ0, double boxZ0) {
  // setOriginsFromGamma / gammaBetaFromGamma, boostframe.cpp:42-72
  if (!(gamma >= 1.0) || !std::isfinite(gamma)) throw std::invalid_argument("Lorentz gamma must be finite and >= 1");
  const double gb = std::sqrt(gamma - 1.0) * std::sqrt(gamma + 1.0);
  if (!std::isfinite(labT0) || !std::isfinite(labZ0) || !std::isfinite(boxZ0))
    throw std::invalid_argument("Lorentz origins must be finite");
  return FrameDev{gamma, gb / gamma, gb, labT0, labZ0, boxZ0};
}

__device__ __forceinline__ double positiveLightFront(const double u[3], unsigned& err) {
  const double g = draw::gammaOf(u);
  const double mt2 = fma(u[0], u[0], fma(u[1], u[1], 1.0));
  if (!(mt2 > 0.0) || !isfinite(mt2)) err |= kErrTransverseMass;
  return u[2] >= 0.0 ? g + u[2] : mt2 / (g - u[2]);
}
__device__ __forceinline__ void fromPositiveLightFront(const double in[3], double positive, double out[3],
                                                       unsigned& err) {
  const double mt2 = fma(in[0], in[0], fma(in[1], in[1], 1.0));
  if (!(positive > 0.0) || !isfinite(positive)) err |= kErrLightFront;
  const double negative = mt2 / positive;
  out[0] = in[0];
  out[1] = in[1];
  out[2] = 0.5 * (positive - negative);
  if (!isfinite(out[2])) err |= kErrLongitudinal;
}
// properVelocityLabToBox / properVelocityBoxToLab, boostframe.cpp:147-164
__device__ __forceinline__ void labToBoxU(const FrameDev& f, const double in[3], double out[3], unsigned& err) {
  fromPositiveLightFront(in, positiveLightFront(in, err) / (f.gamma + f.gammaBeta), out, err);
}
__device__ __forceinline__ void boxToLabU(const FrameDev& f, const double in[3], double out[3], unsigned& err) {
  fromPositiveLightFront(in, positiveLightFront(in, err) * (f.gamma + f.gammaBeta), out, err);
}
// heldLongitudinalBetaZ + longitudinalOnlyBetaZ, boostframe.cpp:90-107
__device__ __forceinline__ double heldBetaZ(const FrameDev& f, const double uBox[3], unsigned& err) {
  double lab[3], longitudinal[3] = {0.0, 0.0, 0.0}, t[3];
  boxToLabU(f, uBox, lab, err);
  longitudinal[2] = lab[2];
  labToBoxU(f, longitudinal, t, err);
  return t[2] / hypot(1.0, t[2]);
}

__device__ __forceinline__ void loadU(const Bunch& B, long long p, double u[3]) {
  for (int a = 0; a < 3; ++a) u[a] = B.u[a][p];
}

// One atomic per block for each of K order-preserving slots, a minimum or
// (isMax[c]) a maximum; v[c] is this lane's value for slot c.
template <int K>
__device__ void blockExtrema(const double v[K], const bool isMax[K], unsigned long long* slots) {
  __shared__ unsigned long long sh[K][kThreads];
  for (int c = 0; c < K; ++c) sh[c][threadIdx.x] = orderedBits(v[c]);
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s)
      for (int c = 0; c < K; ++c) {
        const unsigned long long a = sh[c][threadIdx.x], b = sh[c][threadIdx.x + s];
        sh[c][threadIdx.x] = isMax[c] ? (a > b ? a : b) : (a < b ? a : b);
      }
    __syncthreads();
  }
  if (threadIdx.x < static_cast<unsigned>(K)) {
    const int c = threadIdx.x;
    if (isMax[c]) atomicMax(slots + c, sh[c][0]);
    else atomicMin(slots + c, sh[c][0]);
  }
}

__device__ __forceinline__ void blockError(unsigned e, unsigned* err) {
  // one atomic per wavefront with a nonzero word
  for (int off = 32; off > 0; off >>= 1) e |= __shfl_xor(e, off);
  if (e && (threadIdx.x & 63) == 0) atomicOr(err, e);
}

// Plane-arrival events: plane x, y and the arrival time offset (axis 2, s)
// from the generated-gaussian draws, the lab proper velocity, unit weight.
struct EventArgs {
  double center[3], sigma[3], meanU[3], sigmaU[3];
  double charge, mass;
  unsigned long long seed, first;
};
__global__ void __launch_bounds__(kThreads) kPlaneEvents(EventArgs E, Bunch B, double* q, double* m,
                                                         uint64_t* id) {
  const long long k = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x;
  if (k >= B.n) return;
  const unsigned long long g = E.first + static_cast<unsigned long long>(k);
  for (unsigned a = 0; a < 3; ++a) {
    B.x[a][k] = E.center[a] + E.sigma[a] * draw::gaussianDraw(E.seed, g, a);
    B.u[a][k] = E.meanU[a] + E.sigmaU[a] * draw::gaussianDraw(E.seed ^ 0xd6e8feb86659fd93ull, g, a);
  }
  B.w[k] = 1.0;
  B.beta[k] = 0.0;
  q[k] = E.charge;
  m[k] = E.mass;
  id[k] = g + 1ull;
}

// projectLabPlaneEventsToSnapshot's first loop (eb_bunch.cpp:163-186):
// sum w, sum w beta_z, sum w beta_z tau (double-double) and max tau.
__global__ void __launch_bounds__(kThreads) kProjectSums(Bunch B, DD* partial, unsigned long long* latest,
                                                         unsigned* err) {
  DD v[3];
  double latestLane = -INFINITY;
  unsigned e = 0;
  for (long long p = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x; p < B.n;
       p += static_cast<long long>(gridDim.x) * kThreads) {
    double u[3];
    loadU(B, p, u);
    const double tau = B.x[2][p], w = B.w[p];
    const double gamma = sqrt(1.0 + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
    const double betaZ = u[2] / gamma;
    if (!(w > 0.0) || !isfinite(w) || !(betaZ > 0.0) || !isfinite(betaZ) || !isfinite(tau)) e |= kErrEventInput;
    v[0] = ddAdd(v[0], w);
    v[1] = ddAddProduct(v[1], w, betaZ);
    // w beta_z tau with w beta_z = pr + er exactly
    const double pr = w * betaZ, er = fma(w, betaZ, -pr);
    v[2] = ddAddProduct(ddAddProduct(v[2], pr, tau), er, tau);
    latestLane = fmax(latestLane, tau);
  }
  blockDD<3>(v, partial);
  const bool isMax[1] = {true};
  const double ext[1] = {latestLane};
  blockExtrema<1>(ext, isMax, latest);
  blockError(e, err);
}

struct ProjectArgs {
  double referenceTime, tolerance, inputPlane, reference, offset;
  int hold;
};
// eb_bunch.cpp:236-257: forward distance, its extrema and the snapshot.
__global__ void __launch_bounds__(kThreads) kProject(Bunch B, ProjectArgs A, unsigned long long* extrema) {
  double lo = INFINITY, hi = -INFINITY;
  for (long long p = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x; p < B.n;
       p += static_cast<long long>(gridDim.x) * kThreads) {
    double u[3];
    loadU(B, p, u);
    const double tau = B.x[2][p];
    const double gamma = sqrt(1.0 + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
    const double betaZ = u[2] / gamma;
    const double deltaTime = A.referenceTime - tau;
    double distance = betaZ * kC * deltaTime;
    if (distance < 0.0 && distance >= -A.tolerance) distance = 0.0;
    lo = fmin(lo, distance);
    hi = fmax(hi, distance);
    if (!A.hold) {
      B.x[0][p] += u[0] / u[2] * distance;
      B.x[1][p] += u[1] / u[2] * distance;
    }
    B.x[2][p] = A.inputPlane + distance - A.reference + A.offset;
  }
  const bool isMax[2] = {false, true};
  const double ext[2] = {lo, hi};
  blockExtrema<2>(ext, isMax, extrema);
}

// placeRelativeLabSnapshot (eb_bunch.cpp:313-321): tail and head of z.
__global__ void __launch_bounds__(kThreads) kPlaceExtrema(Bunch B, unsigned long long* extrema, unsigned* err) {
  double lo = INFINITY, hi = -INFINITY;
  unsigned e = 0;
  for (long long p = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x; p < B.n;
       p += static_cast<long long>(gridDim.x) * kThreads) {
    const double z = B.x[2][p];
    if (!isfinite(z)) e |= kErrPositionFinite;
    lo = fmin(lo, z);
    hi = fmax(hi, z);
  }
  const bool isMax[2] = {false, true};
  const double ext[2] = {lo, hi};
  blockExtrema<2>(ext, isMax, extrema);
  blockError(e, err);
}
__global__ void __launch_bounds__(kThreads) kPlaceShift(Bunch B, double reference) {
  for (long long p = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x; p < B.n;
       p += static_cast<long long>(gridDim.x) * kThreads)
    B.x[2][p] += reference;
}

struct OriginArgs {
  FrameDev probe;
  double inputPlane, activationPlane, dz;
  int hold;
};
// simulation.cpp:2132-2173: each particle's earliest crossing time.
__global__ void __launch_bounds__(kThreads) kFrozenOrigin(Bunch B, OriginArgs A, unsigned long long* earliest,
                                                          unsigned* err) {
  double lo = INFINITY;
  unsigned e = 0;
  for (long long p = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x; p < B.n;
       p += static_cast<long long>(gridDim.x) * kThreads) {
    double u[3];
    loadU(B, p, u);
    const double gamma = draw::gammaOf(u);
    const double betaZ = u[2] / gamma;
    const double snapshotZ = B.x[2][p];
    if (A.hold) {
      double uBox[3];
      labToBoxU(A.probe, u, uBox, e);
      const double betaHeld = heldBetaZ(A.probe, uBox, e);
      const double labAdvance = A.probe.beta + betaHeld;
      if (!(labAdvance > 0.0) || !isfinite(labAdvance)) {
        e |= kErrHeldAdvance;
        continue;
      }
      const double inputTime = (A.inputPlane - snapshotZ) / (betaZ * kC);
      const double inputToActivation = A.inputPlane - A.activationPlane;
      const double crossing =
          inputTime - inputToActivation * (1.0 + A.probe.beta * betaHeld) / (kC * labAdvance);
      const double margin = 2.0 * A.dz / (A.probe.gamma * kC * labAdvance);
      lo = fmin(lo, crossing - margin);
    } else {
      lo = fmin(lo, (A.activationPlane - snapshotZ) / (betaZ * kC));
    }
  }
  const bool isMax[1] = {false};
  const double ext[1] = {lo};
  blockExtrema<1>(ext, isMax, earliest);
  blockError(e, err);
}

struct BoostArgs {
  FrameDev f;
  double snapshotTime, inputPlane;
  int hold;
};
// boostLabSnapshotToBoxTimeZero's loop, eb_bunch.cpp:372-472; report slots:
// earliest (min), latest (max), |drift| (max), lab gamma (min, max),
// momentum and gamma round-trip errors (max).
__global__ void __launch_bounds__(kThreads) kBoost(Bunch B, BoostArgs A, unsigned long long* report, unsigned* err) {
  double ext[7] = {INFINITY, -INFINITY, 0.0, INFINITY, 1.0, 0.0, 0.0};
  unsigned e = 0;
  const FrameDev& f = A.f;
  for (long long p = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x; p < B.n;
       p += static_cast<long long>(gridDim.x) * kThreads) {
    double uLab[3], pos[3];
    loadU(B, p, uLab);
    for (int a = 0; a < 3; ++a) pos[a] = B.x[a][p];
    const double gammaParticle = draw::gammaOf(uLab);
    double betaParticle[3];
    for (int a = 0; a < 3; ++a) betaParticle[a] = (1.0 / gammaParticle) * uLab[a];
    double eventTime = 0.0, drift = 0.0, eventZBox = 0.0;
    double eventPos[3] = {pos[0], pos[1], pos[2]};
    if (A.hold) {
      if (!(betaParticle[2] > 0.0)) {
        e |= kErrHeldVelocity;
        continue;
      }
      const double inputTime = A.snapshotTime + (A.inputPlane - pos[2]) / (kC * betaParticle[2]);
      // labToBox, boostframe.cpp:172-180
      const double time = inputTime - f.labT0, position = A.inputPlane - f.labZ0;
      const double inputTimeBox = fma(-f.gammaBeta / kC, position, f.gamma * time);
      const double inputZBox = f.boxZ0 + fma(-f.gammaBeta * kC, time, f.gamma * position);
      double uBox[3];
      labToBoxU(f, uLab, uBox, e);
      const double betaHeld = heldBetaZ(f, uBox, e);
      if (!(f.beta + betaHeld > 0.0)) {
        e |= kErrHeldLab;
        continue;
      }
      eventZBox = fma(-kC * betaHeld, inputTimeBox, inputZBox);
      eventTime = f.labT0 + fma(f.gammaBeta / kC, eventZBox - f.boxZ0, f.gamma * 0.0);
      drift = eventTime - A.snapshotTime;
    } else {
      const double denominator = 1.0 - f.beta * betaParticle[2];
      if (!(denominator > 0.0) || !isfinite(denominator)) {
        e |= kErrSimultaneity;
        continue;
      }
      eventTime = (f.labT0 + f.beta * (pos[2] - f.labZ0) / kC - f.beta * betaParticle[2] * A.snapshotTime) /
                  denominator;
      drift = eventTime - A.snapshotTime;
      for (int a = 0; a < 3; ++a) eventPos[a] += kC * drift * betaParticle[a];
      const double time = eventTime - f.labT0, position = eventPos[2] - f.labZ0;
      const double eventTimeBox = fma(-f.gammaBeta / kC, position, f.gamma * time);
      eventZBox = f.boxZ0 + fma(-f.gammaBeta * kC, time, f.gamma * position);
      const double timeTolerance = 128.0 * kEps * fmax(1.0, fabs(eventTime));
      if (fabs(eventTimeBox) > timeTolerance) e |= kErrBoxTimeZero;
    }
    if (!A.hold) {
      B.x[0][p] = eventPos[0];
      B.x[1][p] = eventPos[1];
    }
    B.x[2][p] = eventZBox;
    double uBox[3], back[3];
    labToBoxU(f, uLab, uBox, e);
    for (int a = 0; a < 3; ++a) B.u[a][p] = uBox[a];
    boxToLabU(f, uBox, back, e);
    double delta[3];
    for (int a = 0; a < 3; ++a) delta[a] = back[a] - uLab[a];
    const double momentumError = sqrt(delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2]) /
                                 fmax(1.0, sqrt(uLab[0] * uLab[0] + uLab[1] * uLab[1] + uLab[2] * uLab[2]));
    const double backGamma = draw::gammaOf(back);
    ext[0] = fmin(ext[0], eventTime);
    ext[1] = fmax(ext[1], eventTime);
    ext[2] = fmax(ext[2], fabs(drift));
    ext[3] = fmin(ext[3], gammaParticle);
    ext[4] = fmax(ext[4], gammaParticle);
    ext[5] = fmax(ext[5], momentumError);
    ext[6] = fmax(ext[6], fabs(backGamma - gammaParticle) / gammaParticle);
  }
  const bool isMax[7] = {false, true, true, false, true, true, true};
  blockExtrema<7>(ext, isMax, report);
  blockError(e, err);
}

// simulation.cpp:2228-2239: the held speed (transverse hold) and whether a
// particle sits at or past the activation plane at box time zero.
__global__ void __launch_bounds__(kThreads) kFreeze(Bunch B, FrameDev f, double activationPlane, int hold,
                                                    unsigned* err) {
  unsigned e = 0;
  for (long long p = blockIdx.x * static_cast<long long>(kThreads) + threadIdx.x; p < B.n;
       p += static_cast<long long>(gridDim.x) * kThreads) {
    double u[3];
    loadU(B, p, u);
    B.beta[p] = hold ? heldBetaZ(f, u, e) : 0.0;
    const double labZ = f.labZ0 + fma(f.gammaBeta * kC, 0.0, f.gamma * (B.x[2][p] - f.boxZ0));
    if (labZ >= activationPlane) e |= kErrAlreadyPast;
  }
  blockError(e, err);
}

// Start values of the extrema slots and a cleared error word (a fill
// kernel, so that no host buffer has to outlive the call).
struct SlotStart {
  unsigned long long v[8];
};
__global__ void kClearSlots(SlotStart S, int K, unsigned long long* slots, unsigned* err) {
  const int c = threadIdx.x;
  if (c < K) slots[c] = S.v[c];
  if (c == 0) *err = 0u;
}

// The reference's exception for each raised bit, in the order its checks
// run for one particle.
void raiseRouteErrors(unsigned e) {
  if (!e) return;
  if (e & kErrEventInput)
    throw std::invalid_argument(
        "Lab-plane event input requires finite positive weights, finite arrival times, and positive longitudinal "
        "velocity");
  if (e & kErrPositionFinite) throw std::invalid_argument("Bunch position must be finite");
  if (e & kErrHeldVelocity)
    throw std::invalid_argument("Held-transverse input requires positive laboratory z velocity");
  if (e & kErrTransverseMass)
    throw std::overflow_error("Cannot form a finite transverse mass for Lorentz transformation");
  if (e & kErrLightFront) throw std::overflow_error("Lorentz transformation produced an invalid light-front momentum");
  if (e & kErrLongitudinal)
    throw std::overflow_error("Lorentz transformation produced a non-finite longitudinal momentum");
  if (e & kErrHeldAdvance)
    throw std::invalid_argument("Held-transverse pseudo-trajectory must advance in laboratory z for every particle");
  if (e & kErrHeldLab)
    throw std::invalid_argument("Held-transverse pseudo-trajectory does not advance in laboratory z");
  if (e & kErrSimultaneity)
    throw std::runtime_error("Particle does not intersect the requested boosted simultaneity plane");
  if (e & kErrBoxTimeZero) throw std::runtime_error("Free-drift bunch transform did not reach box time zero");
  if (e & kErrAlreadyPast)
    throw std::runtime_error(
        "Frozen lab-plane injection requires every particle upstream of the input plane at box time zero");
  throw std::runtime_error("Unknown input-route error");
}

}  // namespace
}  // namespace aprl_gpu

using namespace aprl_gpu;

// A device-resident macro-particle set in the laboratory coordinates of the
// input route (the reference's particles_ vector before the boost): SoA
// x y z, ux uy uz, charge, mass, weight, id and the held speed.
struct aprl_bunch {
  std::size_t capacity = 0, count = 0;
  DeviceArray<double> x[3], u[3], q, m, w, beta;
  DeviceArray<uint64_t> id;
  // readback words: extrema slots (7), error word, sums (3 double-doubles)
  DeviceArray<unsigned long long> slots;
  DeviceArray<unsigned> err;
  DeviceArray<ddsum::DD> partial, sums;
  Stream stream;

  Bunch args() {
    Bunch B{};
    for (int a = 0; a < 3; ++a) {
      B.x[a] = x[a].get();
      B.u[a] = u[a].get();
    }
    B.w = w.get();
    B.beta = beta.get();
    B.n = static_cast<long long>(count);
    return B;
  }
  void reserve(std::size_t n) {
    if (n <= capacity) return;
    const std::size_t c = std::max<std::size_t>(n, 1);
    for (int a = 0; a < 3; ++a) {
      x[a].resize(c);
      u[a].resize(c);
    }
    q.resize(c);
    m.resize(c);
    w.resize(c);
    beta.resize(c);
    id.resize(c);
    capacity = c;
    count = 0;
  }
  unsigned blocks() const {
    return std::max(1u, std::min<unsigned>(kMaxBlocks, ceilDiv(count ? count : 1, kThreads)));
  }
  // clear the K extrema slots (min slots first where isMax is false) and the
  // error word with small fill kernels
  void clearSlots(const bool* isMax, int K) {
    if (!slots.size()) {
      slots.resize(8);
      err.resize(1);
    }
    SlotStart S{};
    for (int c = 0; c < K; ++c) S.v[c] = isMax[c] ? kOrderedMaxStart : kOrderedMinStart;
    kClearSlots<<<1, 64, 0, stream.get()>>>(S, K, slots.get(), err.get());
    launchCheck("clear slots");
  }
  // read back K extrema and the error word
  unsigned readSlots(double* out, int K) {
    unsigned long long raw[8];
    unsigned e = 0;
    hipCheck(hipMemcpyAsync(raw, slots.get(), K * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream.get()),
             "hipMemcpyAsync");
    hipCheck(hipMemcpyAsync(&e, err.get(), sizeof e, hipMemcpyDeviceToHost, stream.get()), "hipMemcpyAsync");
    hipCheck(hipStreamSynchronize(stream.get()), "hipStreamSynchronize");
    for (int c = 0; c < K; ++c) out[c] = fromOrderedBits(raw[c]);
    return e;
  }
};

namespace {
void requireBunch(const aprl_bunch* b) {
  if (!b) throw std::invalid_argument("Null bunch");
}
long double ldOf(const ddsum::DD& d) { return static_cast<long double>(d.hi) + static_cast<long double>(d.lo); }
}  // namespace

extern "C" {

aprl_status aprl_bunch_create(size_t capacity, aprl_bunch** out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("Null output handle");
    auto b = std::make_unique<aprl_bunch>();
    b->reserve(capacity);
    *out = b.release();
  });
}

void aprl_bunch_destroy(aprl_bunch* bunch) { delete bunch; }

size_t aprl_bunch_count(const aprl_bunch* bunch) { return bunch ? bunch->count : 0; }

aprl_status aprl_bunch_upload(aprl_bunch* bunch, size_t n, const double* position3, const double* u3,
                              const double* charge, const double* mass, const double* weight, const uint64_t* id) {
  return guarded([&] {
    requireBunch(bunch);
    if (n && (!position3 || !u3 || !charge || !mass || !id)) throw std::invalid_argument("Null bunch arrays");
    bunch->reserve(n);
    std::vector<double> col(n);
    hipStream_t st = bunch->stream.get();
    hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
    auto put = [&](DeviceArray<double>& dst, const std::vector<double>& src) {
      if (n) hipCheck(hipMemcpy(dst.get(), src.data(), n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    };
    for (int a = 0; a < 3; ++a) {
      for (size_t p = 0; p < n; ++p) col[p] = position3[3 * p + a];
      put(bunch->x[a], col);
      for (size_t p = 0; p < n; ++p) col[p] = u3[3 * p + a];
      put(bunch->u[a], col);
    }
    put(bunch->q, std::vector<double>(charge, charge + n));
    put(bunch->m, std::vector<double>(mass, mass + n));
    put(bunch->w, weight ? std::vector<double>(weight, weight + n) : std::vector<double>(n, 1.0));
    put(bunch->beta, std::vector<double>(n, 0.0));
    if (n) hipCheck(hipMemcpy(bunch->id.get(), id, n * sizeof(uint64_t), hipMemcpyHostToDevice), "hipMemcpy");
    bunch->count = n;
  });
}

aprl_status aprl_bunch_download(aprl_bunch* bunch, double* position3, double* u3, double* charge, double* mass,
                                double* weight, uint64_t* id, double* held_beta) {
  return guarded([&] {
    requireBunch(bunch);
    const size_t n = bunch->count;
    hipCheck(hipStreamSynchronize(bunch->stream.get()), "hipStreamSynchronize");
    std::vector<double> col(n);
    auto get = [&](const DeviceArray<double>& src) {
      if (n) hipCheck(hipMemcpy(col.data(), src.get(), n * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    };
    for (int a = 0; a < 3; ++a) {
      if (position3) {
        get(bunch->x[a]);
        for (size_t p = 0; p < n; ++p) position3[3 * p + a] = col[p];
      }
      if (u3) {
        get(bunch->u[a]);
        for (size_t p = 0; p < n; ++p) u3[3 * p + a] = col[p];
      }
    }
    auto copyOut = [&](double* dst, const DeviceArray<double>& src) {
      if (dst && n) hipCheck(hipMemcpy(dst, src.get(), n * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    };
    copyOut(charge, bunch->q);
    copyOut(mass, bunch->m);
    copyOut(weight, bunch->w);
    copyOut(held_beta, bunch->beta);
    if (id && n) hipCheck(hipMemcpy(id, bunch->id.get(), n * sizeof(uint64_t), hipMemcpyDeviceToHost), "hipMemcpy");
  });
}

aprl_status aprl_bunch_generate_plane_events(aprl_bunch* bunch, const aprl_gaussian_beam* beam, size_t count) {
  return guarded([&] {
    requireBunch(bunch);
    if (!beam) throw std::invalid_argument("Null beam");
    if (!(beam->total > 0)) throw std::invalid_argument("generated-gaussian requires macroparticles > 0");
    if (beam->first + count > beam->total) throw std::invalid_argument("Beam slice exceeds the beam's macroparticles");
    if (!(beam->electrons > 0.0) || !std::isfinite(beam->electrons))
      throw std::invalid_argument("Beam electrons must be positive and finite");
    if (beam->boost_gamma != 1.0) throw std::invalid_argument("Plane-arrival events are laboratory records (boost 1)");
    EventArgs E{};
    for (int a = 0; a < 3; ++a) {
      if (!std::isfinite(beam->center[a]) || !std::isfinite(beam->sigma_position[a]) ||
          !std::isfinite(beam->mean_u[a]) || !std::isfinite(beam->sigma_u[a]))
        throw std::invalid_argument("Beam parameters must be finite");
      E.center[a] = beam->center[a];
      E.sigma[a] = beam->sigma_position[a];
      E.meanU[a] = beam->mean_u[a];
      E.sigmaU[a] = beam->sigma_u[a];
    }
    // particle_initializer.cpp:69-70, 94-95
    const double macro = beam->electrons / static_cast<double>(beam->total);
    E.charge = -kQe * macro;
    E.mass = kMe * macro;
    E.seed = beam->seed;
    E.first = beam->first;
    bunch->reserve(count);
    bunch->count = count;
    hipStream_t st = bunch->stream.get();
    if (count)
      kPlaneEvents<<<ceilDiv(count, kThreads), kThreads, 0, st>>>(E, bunch->args(), bunch->q.get(), bunch->m.get(),
                                                                    bunch->id.get());
    launchCheck("plane events");
    hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
  });
}

aprl_status aprl_bunch_project_lab_plane_events(aprl_bunch* bunch, double input_plane_lab,
                                                double reference_position_lab, double longitudinal_offset_lab,
                                                int hold_transverse, aprl_lab_plane_projection* report) {
  return guarded([&] {
    requireBunch(bunch);
    if (!std::isfinite(input_plane_lab) || !std::isfinite(reference_position_lab) ||
        !std::isfinite(longitudinal_offset_lab))
      throw std::invalid_argument("Lab-plane event projection coordinates must be finite");
    if (!bunch->count) throw std::invalid_argument("Cannot project an empty particle event set");
    hipStream_t st = bunch->stream.get();
    const unsigned blocks = bunch->blocks();
    if (bunch->partial.size() < 3 * static_cast<size_t>(kMaxBlocks)) {
      bunch->partial.resize(3 * kMaxBlocks);
      bunch->sums.resize(3);
    }
    const bool maxOnly[1] = {true};
    bunch->clearSlots(maxOnly, 1);
    kProjectSums<<<blocks, kThreads, 0, st>>>(bunch->args(), bunch->partial.get(), bunch->slots.get(),
                                              bunch->err.get());
    launchCheck("projection sums");
    kFinishDD<<<1, 64, 0, st>>>(bunch->partial.get(), static_cast<int>(blocks), 3, bunch->sums.get());
    launchCheck("projection sums");
    ddsum::DD s[3];
    hipCheck(hipMemcpyAsync(s, bunch->sums.get(), sizeof s, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    double latest = 0.0;
    raiseRouteErrors(bunch->readSlots(&latest, 1));
    // eb_bunch.cpp:192-234 in long double from the compensated sums
    const long double weight = ldOf(s[0]), weightedBeta = ldOf(s[1]), weightedBetaTime = ldOf(s[2]);
    if (!(weight > 0.0L) || !(weightedBeta > 0.0L))
      throw std::invalid_argument("Lab-plane event input has an invalid global weight or longitudinal velocity sum");
    const long double meanBeta = weightedBeta / weight;
    const long double referenceTime =
        (static_cast<long double>(reference_position_lab - input_plane_lab) / static_cast<long double>(kC) +
         weightedBetaTime / weight) /
        meanBeta;
    const double scale = std::max(1.0, std::max(std::abs(input_plane_lab), std::abs(reference_position_lab)));
    const double tolerance = 128.0 * kEps * scale;
    const long double minimumCenter =
        static_cast<long double>(input_plane_lab) +
        static_cast<long double>(kC) * (meanBeta * static_cast<long double>(latest) - weightedBetaTime / weight);
    if (referenceTime < static_cast<long double>(latest) - static_cast<long double>(tolerance / kC)) {
      std::ostringstream m;
      m << std::setprecision(16)
        << "The reconstructed laboratory snapshot would require backward propagation of at least one HDF5 "
           "lab-plane event. Set beam.reference.initial_center_z to at least "
        << static_cast<double>(minimumCenter)
        << " m, move input_plane_z upstream, or verify the input arrival-time values and selected page.";
      throw std::invalid_argument(m.str());
    }
    ProjectArgs A{static_cast<double>(referenceTime), tolerance, input_plane_lab, reference_position_lab,
                  longitudinal_offset_lab, hold_transverse ? 1 : 0};
    const bool minMax[2] = {false, true};
    bunch->clearSlots(minMax, 2);
    kProject<<<blocks, kThreads, 0, st>>>(bunch->args(), A, bunch->slots.get());
    launchCheck("projection");
    double ext[2];
    raiseRouteErrors(bunch->readSlots(ext, 2));
    if (report) {
      report->input_plane_lab = input_plane_lab;
      report->reference_position_lab = reference_position_lab;
      report->minimum_forward_distance = std::max(0.0, ext[0]);
      report->maximum_forward_distance = ext[1];
      report->recommended_minimum_reference_position = static_cast<double>(minimumCenter);
      report->reference_time_offset_lab = static_cast<double>(referenceTime);
      report->mean_longitudinal_beta = static_cast<double>(meanBeta);
      report->particles = bunch->count;
    }
  });
}

aprl_status aprl_bunch_place_relative(aprl_bunch* bunch, int has_entrance, double entrance_lab,
                                      double reference_position_lab, double margin_lab,
                                      aprl_bunch_placement* report) {
  return guarded([&] {
    requireBunch(bunch);
    if ((has_entrance && !std::isfinite(entrance_lab)) || !std::isfinite(reference_position_lab) ||
        !std::isfinite(margin_lab) || margin_lab < 0.0)
      throw std::invalid_argument(
          "Relative bunch placement requires finite lab coordinates and a nonnegative margin");
    if (!bunch->count) throw std::invalid_argument("Cannot place an empty particle bunch");
    hipStream_t st = bunch->stream.get();
    const unsigned blocks = bunch->blocks();
    const bool minMax[2] = {false, true};
    bunch->clearSlots(minMax, 2);
    kPlaceExtrema<<<blocks, kThreads, 0, st>>>(bunch->args(), bunch->slots.get(), bunch->err.get());
    launchCheck("placement");
    double ext[2];
    raiseRouteErrors(bunch->readSlots(ext, 2));
    kPlaceShift<<<blocks, kThreads, 0, st>>>(bunch->args(), reference_position_lab);
    launchCheck("placement");
    hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (report) {
      report->particles = bunch->count;
      report->relative_tail_lab = ext[0];
      report->relative_head_lab = ext[1];
      report->reference_position_lab = reference_position_lab;
      report->head_after_lab = reference_position_lab + ext[1];
      if (has_entrance) {
        report->actual_head_distance = entrance_lab - report->head_after_lab;
        report->recommended_maximum_reference_position = entrance_lab - ext[1] - margin_lab;
      } else {
        report->actual_head_distance = std::numeric_limits<double>::infinity();
        report->recommended_maximum_reference_position = std::numeric_limits<double>::infinity();
      }
    }
  });
}

aprl_status aprl_bunch_frozen_frame_origin(aprl_bunch* bunch, double gamma_boost, double input_plane_lab,
                                           double activation_plane_lab, double dz, int hold_transverse,
                                           aprl_boost* frame) {
  return guarded([&] {
    requireBunch(bunch);
    if (!frame) throw std::invalid_argument("Null frame");
    OriginArgs A{};
    A.probe = frameFrom(gamma_boost, 0.0, activation_plane_lab, 0.0);
    A.inputPlane = input_plane_lab;
    A.activationPlane = activation_plane_lab;
    A.dz = dz;
    A.hold = hold_transverse ? 1 : 0;
    const bool minOnly[1] = {false};
    bunch->clearSlots(minOnly, 1);
    kFrozenOrigin<<<bunch->blocks(), kThreads, 0, bunch->stream.get()>>>(bunch->args(), A, bunch->slots.get(),
                                                                         bunch->err.get());
    launchCheck("frozen frame origin");
    double earliest = std::numeric_limits<double>::infinity();
    raiseRouteErrors(bunch->readSlots(&earliest, 1));
    if (!bunch->count) earliest = std::numeric_limits<double>::infinity();
    frame->gamma = gamma_boost;
    frame->lab_time_origin = hold_transverse ? earliest : earliest - 2.0 * dz / (gamma_boost * kC);
    frame->lab_z_origin = activation_plane_lab;
    frame->box_z_origin = 0.0;
  });
}

aprl_status aprl_bunch_boost_to_box_time_zero(aprl_bunch* bunch, const aprl_boost* frame, double snapshot_time_lab,
                                              int hold_transverse, double input_plane_lab,
                                              aprl_bunch_boost_report* report) {
  return guarded([&] {
    requireBunch(bunch);
    if (!frame) throw std::invalid_argument("Null frame");
    if (!std::isfinite(snapshot_time_lab) || (hold_transverse && !std::isfinite(input_plane_lab)))
      throw std::invalid_argument("Bunch snapshot time and held-mode input plane must be finite");
    BoostArgs A{};
    A.f = frameFrom(frame->gamma, frame->lab_time_origin, frame->lab_z_origin, frame->box_z_origin);
    A.snapshotTime = snapshot_time_lab;
    A.inputPlane = input_plane_lab;
    A.hold = hold_transverse ? 1 : 0;
    const bool slotsMax[7] = {false, true, true, false, true, true, true};
    bunch->clearSlots(slotsMax, 7);
    kBoost<<<bunch->blocks(), kThreads, 0, bunch->stream.get()>>>(bunch->args(), A, bunch->slots.get(),
                                                                  bunch->err.get());
    launchCheck("boost to box time zero");
    double ext[7];
    raiseRouteErrors(bunch->readSlots(ext, 7));
    if (report) {
      // the report's defaults (eb_bunch.cpp:32-39) where no particle set it
      report->earliest_lab_event_time = bunch->count ? ext[0] : 0.0;
      report->latest_lab_event_time = bunch->count ? ext[1] : 0.0;
      report->maximum_absolute_drift_time = std::max(0.0, ext[2]);
      report->minimum_lab_gamma = bunch->count ? ext[3] : std::numeric_limits<double>::infinity();
      report->maximum_lab_gamma = bunch->count ? std::max(1.0, ext[4]) : 1.0;
      report->maximum_relative_momentum_round_trip_error = std::max(0.0, ext[5]);
      report->maximum_relative_gamma_round_trip_error = std::max(0.0, ext[6]);
    }
  });
}

aprl_status aprl_bunch_freeze(aprl_bunch* bunch, const aprl_boost* frame, double activation_plane_lab,
                              int hold_transverse) {
  return guarded([&] {
    requireBunch(bunch);
    if (!frame) throw std::invalid_argument("Null frame");
    const FrameDev f = frameFrom(frame->gamma, frame->lab_time_origin, frame->lab_z_origin, frame->box_z_origin);
    const bool none[1] = {false};
    bunch->clearSlots(none, 1);
    kFreeze<<<bunch->blocks(), kThreads, 0, bunch->stream.get()>>>(bunch->args(), f, activation_plane_lab,
                                                                   hold_transverse ? 1 : 0, bunch->err.get());
    launchCheck("freeze");
    double unused = 0.0;
    raiseRouteErrors(bunch->readSlots(&unused, 1));
  });
}

}  // extern "C"
