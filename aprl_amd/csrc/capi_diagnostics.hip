// C-ABI entry points of the FP64 diagnostics (declared in include/aprl_gpu.h),
// for both forms of the grid: the FP32-split form's fields, currents and
// positions are widened and every sum runs in FP64.
#include <cmath>

#include "diagnostics.hpp"
#include "handles.hpp"

using namespace aprl_gpu;

namespace {
void checkFields(aprl_fields* f) {
  if (!f) throw std::invalid_argument("Null field grid handle");
}
void checkCoupled(aprl_coupling* c, aprl_particles* p) {
  if (!c || !p) throw std::invalid_argument("Null coupling or particle store");
  if (p->fields == nullptr || p->fields != c->fields)
    throw std::invalid_argument("Particle store and coupling belong to different field grids");
}
const double* const* positions(ParticleStore& P, bool start, const double* out[3]) {
  for (int a = 0; a < 3; ++a) out[a] = start ? P.end[a].get() : P.s[a].get();
  return out;
}
// Split-form positions: start of the last step (end arrays) or current.
void splitPositions(SplitStore& P, bool start, const int* cell[3], const float* off[3]) {
  for (int a = 0; a < 3; ++a) {
    cell[a] = start ? P.endCell[a].get() : P.cell[a].get();
    off[a] = start ? P.end[a].get() : P.s[a].get();
  }
}
// Charge factors of a store at the start (mode 2) or end (mode 1) of its
// last step: removed entries carry no charge and carriers their weight.
template <class Store, class W>
ChargeWeights weightsOf(Store& P, int mode, const W* w, const W* wStart) {
  ChargeWeights c;
  c.mode = mode;
  c.state = P.state.get();
  c.evt = P.evt.get();
  const W* src = mode == 2 ? wStart : w;
  if constexpr (sizeof(W) == sizeof(double)) c.wd = reinterpret_cast<const double*>(src);
  else c.wf = reinterpret_cast<const float*>(src);
  return c;
}
// Lattice cell of a grid origin of the split form.
void latticeOrigin(const FieldState32& F, const double origin[3], int oc[3]) {
  const double h[3] = {F.g.dx, F.g.dy, F.g.dz};
  for (int a = 0; a < 3; ++a) {
    const double r = origin[a] / h[a];
    if (!std::isfinite(r) || std::fabs(r) >= 1.0e9 || std::fabs(r - std::nearbyint(r)) > 1e-6)
      throw std::invalid_argument("An FP32-split grid origin must lie on the cell lattice (origin = integer * spacing)");
    oc[a] = static_cast<int>(std::nearbyint(r));
  }
}
}  // namespace

namespace {
// Continuity of the last coupled step of this store in the workspace ws
// with the charge densities rho0 / rho1 (one double per vertex each).
void continuityOf(aprl_coupling* coupling, aprl_particles* particles, const RhoWorkspace& ws, double* rho0,
                  double* rho1, double* out) {
  CouplingState& C = coupling->c;
  if (coupling->fields->fp32) {
    SplitStore& P = particles->sp;
    if (!C.lastStep.valid || C.lastStep.store != &P || C.lastStep.generation != P.generation)
      throw std::logic_error("Continuity needs the particle store of the last coupled step");
    FieldState32& F = *C.fields32;
    int oc[3];
    latticeOrigin(F, C.lastStep.origin, oc);
    const int* cell[3];
    const float* off[3];
    const int n = static_cast<int>(P.count);
    splitPositions(P, true, cell, off);
    chargeDensitySplit(F, C.shape, cell, off, P.s[6].get(), n, oc, weightsOf(P, 2, P.s[8].get(), P.wend.get()), ws,
                       rho0);
    splitPositions(P, false, cell, off);
    chargeDensitySplit(F, C.shape, cell, off, P.s[6].get(), n, oc, weightsOf(P, 1, P.s[8].get(), P.wend.get()), ws,
                       rho1);
    continuityResidual(F, rho0, rho1, ws, out);
    return;
  }
  ParticleStore& P = particles->p;
  if (!C.lastStep.valid || C.lastStep.store != &P || C.lastStep.generation != P.generation)
    throw std::logic_error("Continuity needs the particle store of the last coupled step");
  FieldState& F = *C.fields;
  const double* x[3];
  const int n = static_cast<int>(P.count);
  chargeDensity(F, C.shape, positions(P, true, x), P.s[6].get(), n, C.lastStep.origin,
                weightsOf(P, 2, P.s[8].get(), P.wend.get()), ws, rho0);
  chargeDensity(F, C.shape, positions(P, false, x), P.s[6].get(), n, C.lastStep.origin,
                weightsOf(P, 1, P.s[8].get(), P.wend.get()), ws, rho1);
  continuityResidual(F, rho0, rho1, ws, out);
}

template <class R>
std::size_t vertexCount(const FieldStateT<R>& F) {
  return (F.g.nx + 1) * (F.g.ny + 1) * (F.g.nz + 1);
}
}  // namespace

extern "C" {

aprl_status aprl_fields_interior_energy(aprl_fields* fields, const size_t* lo, const size_t* hi, double* energy) {
  return guarded([&] {
    if (!lo || !hi || !energy) throw std::invalid_argument("Null diagnostic argument");
    checkFields(fields);
    *energy = fields->fp32 ? interiorEnergy(fields->s32, lo, hi) : interiorEnergy(fields->s, lo, hi);
  });
}

aprl_status aprl_fields_face_power(aprl_fields* fields, const size_t* lo, const size_t* hi, double* power) {
  return guarded([&] {
    if (!lo || !hi || !power) throw std::invalid_argument("Null diagnostic argument");
    checkFields(fields);
    if (fields->fp32) facePower(fields->s32, lo, hi, power);
    else facePower(fields->s, lo, hi, power);
  });
}

aprl_status aprl_coupling_particle_energy(aprl_coupling* coupling, aprl_particles* particles, double time,
                                          double* out) {
  return guarded([&] {
    checkCoupled(coupling, particles);
    if (!out) throw std::invalid_argument("Null diagnostic argument");
    if (coupling->fields->fp32) particleEnergySums(coupling->fields->s32, coupling->c, particles->sp, time, out);
    else particleEnergySums(coupling->fields->s, coupling->c, particles->p, time, out);
  });
}


aprl_status aprl_coupling_continuity(aprl_coupling* coupling, aprl_particles* particles, double* out) {
  return guarded([&] {
    checkCoupled(coupling, particles);
    if (!out) throw std::invalid_argument("Null diagnostic argument");
    CouplingState& C = coupling->c;
    const bool split = coupling->fields->fp32;
    const std::size_t nv = split ? vertexCount(*C.fields32) : vertexCount(*C.fields);
    for (auto& r : C.rho)
      if (r.size() != nv) r.resize(nv);
    const RhoWorkspace ws = split ? diagnosticWorkspace(*C.fields32) : diagnosticWorkspace(*C.fields);
    continuityOf(coupling, particles, ws, C.rho[0].get(), C.rho[1].get(), out);
  });
}

aprl_status aprl_coupling_continuity_noalloc(aprl_coupling* coupling, aprl_particles* particles, double* out) {
  return guarded([&] {
    checkCoupled(coupling, particles);
    if (!out) throw std::invalid_argument("Null diagnostic argument");
    CouplingState& C = coupling->c;
    const bool split = coupling->fields->fp32;
    double* rho0 = nullptr;
    double* rho1 = nullptr;
    const RhoWorkspace ws = split ? noAllocWorkspace(*C.fields32, vertexCount(*C.fields32), &rho0, &rho1)
                                  : noAllocWorkspace(*C.fields, vertexCount(*C.fields), &rho0, &rho1);
    continuityOf(coupling, particles, ws, rho0, rho1, out);
  });
}

aprl_status aprl_coupling_gauss(aprl_coupling* coupling, aprl_particles* particles, const double* origin,
                                const size_t* lo, const size_t* hi, double* out) {
  return guarded([&] {
    checkCoupled(coupling, particles);
    if (!origin || !lo || !hi || !out) throw std::invalid_argument("Null diagnostic argument");
    CouplingState& C = coupling->c;
    if (coupling->fields->fp32) {
      SplitStore& P = particles->sp;
      FieldState32& F = *C.fields32;
      if (C.rho[0].size() != vertexCount(F)) C.rho[0].resize(vertexCount(F));
      int oc[3];
      latticeOrigin(F, origin, oc);
      const int* cell[3];
      const float* off[3];
      splitPositions(P, false, cell, off);
      chargeDensitySplit(F, C.shape, cell, off, P.s[6].get(), static_cast<int>(P.count), oc,
                         weightsOf(P, 1, P.s[8].get(), P.wend.get()), diagnosticWorkspace(F), C.rho[0].get());
      gaussResidual(F, C.rho[0], lo, hi, out);
      return;
    }
    ParticleStore& P = particles->p;
    FieldState& F = *C.fields;
    if (C.rho[0].size() != vertexCount(F)) C.rho[0].resize(vertexCount(F));
    const double* x[3];
    chargeDensity(F, C.shape, positions(P, false, x), P.s[6].get(), static_cast<int>(P.count), origin,
                  weightsOf(P, 1, P.s[8].get(), P.wend.get()), diagnosticWorkspace(F), C.rho[0].get());
    gaussResidual(F, C.rho[0], lo, hi, out);
  });
}

}  // extern "C"
