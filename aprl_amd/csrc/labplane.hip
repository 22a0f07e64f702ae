// Laboratory field planes: LabDetectorManager::sampleFieldPlanes
// (lab_detectors.cpp:2426-2487) for one rank, with samplePlane (2836-2898;
// the simulation passes no external background, simulation.cpp:3291-3293),
// sampleDetectorComponent / interpolationAxis (33-97) and ownerRank
// (2811-2826). A plane at fixed laboratory z is sampled on the device after
// a step, at the box z where the lab plane sits at that box time, once per
// lab-time rhythm: every cell centre (i, j) gets E and B from the Yee
// components by the detector's clamped trilinear stencil, transformed to
// the lab frame (BoostFrameTransform::fieldsBoxToLab, boostframe.cpp:
// 267-285). Samples collect in a device ring and move to the host in one
// copy when fetched.
#include <cmath>
#include <limits>
#include <memory>
#include <vector>

#include "handles.hpp"

struct aprl_lab_plane {
  aprl_coupling* coupling = nullptr;
  double zLab = 0.0, rhythm = 0.0;
  bool sampled = false;
  double nextTimeLab = 0.0;
  std::size_t capacity = 0, points = 0;
  aprl_gpu::DeviceArray<double> ring;  // capacity x points x 6
  std::vector<double> times;           // lab times of the samples in the ring
};

namespace aprl_gpu {
namespace {

constexpr double kEps = std::numeric_limits<double>::epsilon();

struct Axis {
  int lower;
  double fraction;
};

// interpolationAxis, lab_detectors.cpp:39-58.
__device__ __forceinline__ Axis interpolationAxis(double lattice, int size) {
  if (lattice <= 0.0) return Axis{0, 0.0};
  if (lattice >= static_cast<double>(size - 1)) return Axis{size - 2, 1.0};
  const int lower = static_cast<int>(floor(lattice));
  return Axis{lower, lattice - static_cast<double>(lower)};
}

// sampleDetectorComponent, lab_detectors.cpp:66-97 (every plane local).
template <class R>
__device__ __forceinline__ double sampleComponent(const R* c, const int* n, double lx, double ly, const Axis& z) {
  const Axis x = interpolationAxis(lx, n[0]);
  const Axis y = interpolationAxis(ly, n[1]);
  double value = 0.0;
  for (int cz = 0; cz < 2; ++cz) {
    const double weightZ = cz ? z.fraction : 1.0 - z.fraction;
    if (weightZ == 0.0) continue;
    const size_t k = static_cast<size_t>(z.lower + cz);
    for (int b = 0; b < 2; ++b)
      for (int a = 0; a < 2; ++a) {
        const double weightX = a ? x.fraction : 1.0 - x.fraction;
        const double weightY = b ? y.fraction : 1.0 - y.fraction;
        const double point = static_cast<double>(c[(k * n[1] + (y.lower + b)) * n[0] + (x.lower + a)]);
        value += weightX * weightY * weightZ * point;
      }
  }
  return value;
}

struct PlaneArgs {
  double coordinate, ox, oy, dx, dy;
  double gamma, gammaBeta, c0;
  int nx, ny, nz;
};

// samplePlane, lab_detectors.cpp:2836-2898: one thread per cell centre,
// records j-major, (E_lab, B_lab).
template <class R>
__global__ void kLabPlane(FieldPtrsT<R> F, PlaneArgs A, double* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= A.nx * A.ny) return;
  const int i = t % A.nx, j = t / A.nx;
  const Axis vertexZ = interpolationAxis(A.coordinate, A.nz + 1);
  const Axis halfZ = interpolationAxis(A.coordinate - 0.5, A.nz);
  const double py = A.oy + (static_cast<double>(j) + 0.5) * A.dy;
  const double y = (py - A.oy) / A.dy;
  const double px = A.ox + (static_cast<double>(i) + 0.5) * A.dx;
  const double x = (px - A.ox) / A.dx;
  double e[3], b[3];
  e[0] = sampleComponent(F.e[0], F.n[0], x - 0.5, y, vertexZ);
  e[1] = sampleComponent(F.e[1], F.n[1], x, y - 0.5, vertexZ);
  e[2] = sampleComponent(F.e[2], F.n[2], x, y, halfZ);
  b[0] = sampleComponent(F.b[0], F.n[3], x, y - 0.5, halfZ);
  b[1] = sampleComponent(F.b[1], F.n[4], x - 0.5, y, halfZ);
  b[2] = sampleComponent(F.b[2], F.n[5], x - 0.5, y - 0.5, vertexZ);
  double* r = out + 6 * static_cast<size_t>(t);
  r[0] = fma(A.gammaBeta * A.c0, b[1], A.gamma * e[0]);
  r[1] = fma(-A.gammaBeta * A.c0, b[0], A.gamma * e[1]);
  r[2] = e[2];
  r[3] = fma(-A.gammaBeta / A.c0, e[1], A.gamma * b[0]);
  r[4] = fma(A.gammaBeta / A.c0, e[0], A.gamma * b[1]);
  r[5] = b[2];
}

template <class R>
void launchPlane(FieldStateT<R>& s, const PlaneArgs& A, double* out) {
  const int n = A.nx * A.ny;
  kLabPlane<R><<<ceilDiv(n, 256), 256, 0, s.stream>>>(s.ptrs(), A, out);
  launchCheck("lab field plane");
}

// One sampleFieldPlanes pass for this plane: true when it sampled.
template <class R>
bool samplePlane(aprl_lab_plane& P, FieldStateT<R>& s, const BoostDev& f, double timeBox, const double origin[3]) {
  // BoostFrameTransform::boxZFromLabZAndBoxT (boostframe.cpp:199-203)
  const double boxZ = f.boxZ0 + (P.zLab - f.labZ0) / f.gamma - f.beta * f.c0 * timeBox;
  // ownerRank on one rank: the plane must lie on the grid
  const double coordinate = (boxZ - origin[2]) / s.g.dz;
  const double onGrid = 64.0 * kEps * std::max(1.0, std::abs(coordinate));
  if (coordinate < -onGrid || coordinate > static_cast<double>(s.g.nz) + onGrid) return false;
  // BoostFrameTransform::labTFromLabZBoxT (boostframe.cpp:242-246)
  const double timeLab = f.labT0 + timeBox / f.gamma + f.beta * (P.zLab - f.labZ0) / f.c0;
  const double tolerance = 64.0 * kEps * std::max(P.rhythm, std::max(std::abs(timeLab), std::abs(P.nextTimeLab)));
  if (P.sampled && timeLab < P.nextTimeLab - tolerance) return false;
  if (P.times.size() >= P.capacity)
    throw std::logic_error("Lab field plane buffer is full: fetch its samples before the next one is due");
  PlaneArgs A{};
  A.coordinate = coordinate;
  A.ox = origin[0];
  A.oy = origin[1];
  A.dx = s.g.dx;
  A.dy = s.g.dy;
  A.gamma = f.gamma;
  A.gammaBeta = f.gammaBeta;
  A.c0 = f.c0;
  A.nx = static_cast<int>(s.g.nx);
  A.ny = static_cast<int>(s.g.ny);
  A.nz = static_cast<int>(s.g.nz);
  launchPlane(s, A, P.ring.get() + P.times.size() * P.points * 6);
  P.times.push_back(timeLab);
  if (!P.sampled) {
    P.sampled = true;
    P.nextTimeLab = timeLab + P.rhythm;
  } else {
    do { P.nextTimeLab += P.rhythm; } while (P.nextTimeLab <= timeLab + tolerance);
  }
  return true;
}

}  // namespace
}  // namespace aprl_gpu

using namespace aprl_gpu;

extern "C" {

aprl_status aprl_lab_plane_create(aprl_coupling* coupling, double z_lab, double rhythm, size_t capacity,
                                  aprl_lab_plane** out) {
  return guarded([&] {
    if (!coupling || !out) throw std::invalid_argument("Null coupling or lab plane output");
    if (!std::isfinite(z_lab)) throw std::invalid_argument("Lab field plane z must be finite");
    // config.cpp:1713-1717
    if (!(rhythm > 0.0) || !std::isfinite(rhythm)) throw std::invalid_argument("field detector rhythm must be positive");
    if (capacity == 0) throw std::invalid_argument("Lab field plane capacity must be positive");
    auto p = std::make_unique<aprl_lab_plane>();
    p->coupling = coupling;
    p->zLab = z_lab;
    p->rhythm = rhythm;
    p->capacity = capacity;
    const aprl_fields* f = coupling->fields;
    const auto& g = f->fp32 ? f->s32.g : f->s.g;
    p->points = g.nx * g.ny;
    p->ring.resize(capacity * p->points * 6);
    p->times.reserve(capacity);
    *out = p.release();
  });
}

void aprl_lab_plane_destroy(aprl_lab_plane* plane) { delete plane; }

aprl_status aprl_lab_plane_sample(aprl_lab_plane* plane, double time_box, const double* origin_box, int* taken) {
  return guarded([&] {
    if (!plane || !origin_box) throw std::invalid_argument("Null lab plane or origin");
    aprl_coupling* c = plane->coupling;
    aprl_fields* f = c->fields;
    const bool t = f->fp32 ? samplePlane(*plane, f->s32, c->c.frame, time_box, origin_box)
                           : samplePlane(*plane, f->s, c->c.frame, time_box, origin_box);
    if (taken) *taken = t ? 1 : 0;
  });
}

aprl_status aprl_lab_plane_fetch(aprl_lab_plane* plane, size_t capacity, double* time_lab, double* records,
                                 size_t* count) {
  return guarded([&] {
    if (!plane || !count) throw std::invalid_argument("Null lab plane or count");
    const std::size_t n = plane->times.size();
    if (n > capacity) throw std::invalid_argument("Lab field plane fetch buffer holds fewer samples than are pending");
    if (n && (!time_lab || !records)) throw std::invalid_argument("Null lab field plane fetch buffer");
    aprl_fields* f = plane->coupling->fields;
    hipStream_t st = f->fp32 ? f->s32.stream : f->s.stream;
    if (n) {
      hipCheck(hipMemcpyAsync(records, plane->ring.get(), n * plane->points * 6 * sizeof(double),
                              hipMemcpyDeviceToHost, st),
               "hipMemcpyAsync");
      hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
      for (std::size_t i = 0; i < n; ++i) time_lab[i] = plane->times[i];
    }
    plane->times.clear();
    *count = n;
  });
}

}  // extern "C"
