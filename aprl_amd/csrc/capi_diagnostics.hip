// C-ABI entry points of the FP64 diagnostics (declared in include/aprl_gpu.h).
#include "diagnostics.hpp"
#include "handles.hpp"

using namespace aprl_gpu;

namespace {
FieldState& fieldsOf(aprl_fields* f) {
  if (!f) throw std::invalid_argument("Null field grid handle");
  if (f->fp32) throw std::invalid_argument("This diagnostic entry point takes an APRL_FP64 grid");
  return f->s;
}
void checkCoupled(aprl_coupling* c, aprl_particles* p) {
  if (!c || !p) throw std::invalid_argument("Null coupling or particle store");
  if (p->fields == nullptr || &p->fields->s != c->c.fields)
    throw std::invalid_argument("Particle store and coupling belong to different field grids");
}
const double* const* positions(ParticleStore& P, bool start, const double* out[3]) {
  for (int a = 0; a < 3; ++a) out[a] = start ? P.end[a].get() : P.s[a].get();
  return out;
}
}  // namespace

extern "C" {

aprl_status aprl_fields_interior_energy(aprl_fields* fields, const size_t* lo, const size_t* hi, double* energy) {
  return guarded([&] {
    if (!lo || !hi || !energy) throw std::invalid_argument("Null diagnostic argument");
    *energy = interiorEnergy(fieldsOf(fields), lo, hi);
  });
}

aprl_status aprl_fields_face_power(aprl_fields* fields, const size_t* lo, const size_t* hi, double* power) {
  return guarded([&] {
    if (!lo || !hi || !power) throw std::invalid_argument("Null diagnostic argument");
    facePower(fieldsOf(fields), lo, hi, power);
  });
}

aprl_status aprl_coupling_continuity(aprl_coupling* coupling, aprl_particles* particles, double* out) {
  return guarded([&] {
    checkCoupled(coupling, particles);
    if (!out) throw std::invalid_argument("Null diagnostic argument");
    CouplingState& C = coupling->c;
    ParticleStore& P = particles->p;
    if (!C.lastStep.valid || C.lastStep.store != &P || C.lastStep.generation != P.generation)
      throw std::logic_error("Continuity needs the particle store of the last coupled step");
    FieldState& F = *C.fields;
    const double* x[3];
    const int n = static_cast<int>(P.count);
    chargeDensity(F, C.shape, positions(P, true, x), P.s[6].get(), n, C.lastStep.origin, C.rho[0]);
    chargeDensity(F, C.shape, positions(P, false, x), P.s[6].get(), n, C.lastStep.origin, C.rho[1]);
    continuityResidual(F, C.rho[0], C.rho[1], out);
  });
}

aprl_status aprl_coupling_gauss(aprl_coupling* coupling, aprl_particles* particles, const double* origin,
                                const size_t* lo, const size_t* hi, double* out) {
  return guarded([&] {
    checkCoupled(coupling, particles);
    if (!origin || !lo || !hi || !out) throw std::invalid_argument("Null diagnostic argument");
    CouplingState& C = coupling->c;
    ParticleStore& P = particles->p;
    const double* x[3];
    chargeDensity(*C.fields, C.shape, positions(P, false, x), P.s[6].get(), static_cast<int>(P.count), origin,
                  C.rho[0]);
    gaussResidual(*C.fields, C.rho[0], lo, hi, out);
  });
}

}  // extern "C"
