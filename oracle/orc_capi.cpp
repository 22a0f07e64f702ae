// Oracle C-ABI for the test harness (TEST INFRASTRUCTURE ONLY; never linked
// into the product). Entry points mirror oracle/ref_driver.cpp (prefix ref_)
// so the same Python test drives the restatement and the reference build.
#include <cstring>
#include <memory>
#include <string>

#include "orc_bunch.hpp"
#include "orc_detectors.hpp"
#include "orc_initial_field.hpp"
#include "orc_inlet_reference.hpp"
#include "orc_fields.hpp"
#include "orc_particles.hpp"
#include "orc_pending.hpp"

namespace {
thread_local std::string lastError;
template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::invalid_argument& e) {
    lastError = e.what();
    return 1;
  } catch (const std::out_of_range& e) {
    lastError = e.what();
    return 2;
  } catch (const std::logic_error& e) {
    lastError = e.what();
    return 4;
  } catch (const std::runtime_error& e) {
    lastError = e.what();
    return 3;
  } catch (const std::exception& e) {
    lastError = e.what();
    return 5;
  }
}
orc::Fields* F(void* h) { return static_cast<orc::Fields*>(h); }
// n particles from interleaved position / proper-velocity arrays and weights
std::vector<orc::Particle> particlesOf(std::size_t n, const double* x3, const double* u3, const double* w) {
  std::vector<orc::Particle> P(n);
  for (std::size_t p = 0; p < n; ++p) {
    for (int a = 0; a < 3; ++a) {
      P[p].position[a] = x3[3 * p + a];
      P[p].u[a] = u3 ? u3[3 * p + a] : 0.0;
    }
    P[p].weight = w ? w[p] : 1.0;
  }
  return P;
}
void storeParticles(const std::vector<orc::Particle>& P, double* x3, double* u3) {
  for (std::size_t p = 0; p < P.size(); ++p)
    for (int a = 0; a < 3; ++a) {
      x3[3 * p + a] = P[p].position[a];
      if (u3) u3[3 * p + a] = P[p].u[a];
    }
}
orc::CpmlParams params(const std::size_t cells[3], double order, double R, double kappa, double alpha) {
  orc::CpmlParams p;
  for (int a = 0; a < 3; ++a) p.cells[a] = cells[a];
  p.order = order;
  p.reflection = R;
  p.kappaMax = kappa;
  p.alphaFraction = alpha;
  return p;
}
}  // namespace

extern "C" {

const char* orc_last_error() { return lastError.c_str(); }

// BoostFrameTransform (boostframe.cpp:199-203, 242-246, 267-285) for the
// detector: out = {box z of the lab plane, its lab time, E_lab[3], B_lab[3]}.
int orc_boost_lab(double gamma, double labT0, double labZ0, double boxZ0, double zLab, double tBox,
                  const double* eBox, const double* bBox, double* out8) {
  return guarded([&] {
    const orc::LabFrame f = orc::LabFrame::fromGamma(gamma, labT0, labZ0, boxZ0);
    out8[0] = f.boxZFromLabZAndBoxT(zLab, tBox);
    out8[1] = f.labTFromLabZBoxT(zLab, tBox);
    f.fieldsBoxToLab(eBox, bBox, out8 + 2, out8 + 5);
  });
}

// The laboratory-plane input route (orc_bunch.hpp), in place on interleaved
// arrays. report8 = input plane, reference position, minimum and maximum
// forward distance, recommended minimum reference, reference time offset,
// mean beta_z, particles.
int orc_bunch_project(std::size_t n, double* x3, const double* u3, const double* w, double inputPlane,
                      double reference, double offset, int hold, double* report8) {
  return guarded([&] {
    auto P = particlesOf(n, x3, u3, w);
    const orc::bunch::Projection r = orc::bunch::projectLabPlaneEvents(P, inputPlane, reference, offset, hold != 0);
    storeParticles(P, x3, nullptr);
    const double v[8] = {r.inputPlaneLab, r.referencePositionLab, r.minimumForwardDistance, r.maximumForwardDistance,
                         r.recommendedMinimumReferencePosition, r.referenceTimeOffsetLab, r.meanLongitudinalBeta,
                         static_cast<double>(r.particles)};
    for (int a = 0; a < 8; ++a) report8[a] = v[a];
  });
}
// report7 = relative tail, relative head, reference, head after, actual head
// distance, recommended maximum reference, particles.
int orc_bunch_place(std::size_t n, double* x3, int hasEntrance, double entrance, double reference, double margin,
                    double* report7) {
  return guarded([&] {
    auto P = particlesOf(n, x3, nullptr, nullptr);
    const orc::bunch::Placement r = orc::bunch::placeRelative(P, hasEntrance != 0, entrance, reference, margin);
    storeParticles(P, x3, nullptr);
    const double v[7] = {r.relativeTailLab, r.relativeHeadLab, r.referencePositionLab, r.headAfterLab,
                         r.actualHeadDistance, r.recommendedMaximumReferencePosition,
                         static_cast<double>(r.particles)};
    for (int a = 0; a < 7; ++a) report7[a] = v[a];
  });
}
// out2 = lab time origin, lab z origin of the frozen-injection frame.
int orc_bunch_frozen_origin(std::size_t n, const double* x3, const double* u3, double gammaBoost, double inputPlane,
                            double activationPlane, double dz, int hold, double* out2) {
  return guarded([&] {
    const auto P = particlesOf(n, x3, u3, nullptr);
    orc::bunch::frozenFrameOrigin(P, gammaBoost, inputPlane, activationPlane, dz, hold != 0, out2);
  });
}
// frame4 = gamma, lab time origin, lab z origin, box z origin; report7 =
// earliest and latest lab event time, maximum |drift time|, minimum and
// maximum lab gamma, maximum relative momentum and gamma round-trip errors.
int orc_bunch_boost(std::size_t n, double* x3, double* u3, const double* frame4, double snapshotTime, int hold,
                    double inputPlane, double* report7) {
  return guarded([&] {
    auto P = particlesOf(n, x3, u3, nullptr);
    orc::bunch::Frame f;
    f.set(frame4[0], frame4[1], frame4[2], frame4[3]);
    const orc::bunch::BoostReport r = orc::bunch::boostToBoxTimeZero(P, f, snapshotTime, hold != 0, inputPlane);
    storeParticles(P, x3, u3);
    const double v[7] = {r.earliestLabEventTime, r.latestLabEventTime, r.maximumAbsoluteDriftTime,
                         r.minimumLabGamma, r.maximumLabGamma, r.maximumRelativeMomentumRoundTripError,
                         r.maximumRelativeGammaRoundTripError};
    for (int a = 0; a < 7; ++a) report7[a] = v[a];
  });
}
// beta: the held speed per particle (0 without the transverse hold).
int orc_bunch_freeze(std::size_t n, const double* x3, const double* u3, const double* frame4, double activationPlane,
                     int hold, double* beta) {
  return guarded([&] {
    auto P = particlesOf(n, x3, u3, nullptr);
    orc::bunch::Frame f;
    f.set(frame4[0], frame4[1], frame4[2], frame4[3]);
    orc::bunch::freeze(P, f, activationPlane, hold != 0);
    for (std::size_t p = 0; p < n; ++p) beta[p] = P[p].betaHeld;
  });
}

// The moving window's inlet reference (orc_inlet_reference.hpp).
// spacing4 = dx, dy, dz, dt; shape = -1 CIC, 0 quintic, 1 notch, 2
// linear-binomial3; report4 = minimum mode eigenvalue, minimum pivot,
// source charge, deposited charge.
int orc_inlet_reference_new(const std::size_t* cells3, const double* spacing4, const double* origin3, std::size_t n,
                            const double* x3, const double* q, double beta, const std::size_t* guards3,
                            std::size_t blockLayers, int shape, int xBinomial, void** out, double* report4) {
  return guarded([&] {
    orc::inletref::Config c;
    c.nx = cells3[0];
    c.ny = cells3[1];
    c.nz = cells3[2];
    c.dx = spacing4[0];
    c.dy = spacing4[1];
    c.dz = spacing4[2];
    c.dt = spacing4[3];
    for (int a = 0; a < 3; ++a) {
      c.origin[a] = origin3[a];
      c.guards[a] = guards3[a];
    }
    c.beta = beta;
    c.blockLayers = blockLayers;
    c.shape = shape;
    c.xBinomial = xBinomial != 0;
    auto* r = new orc::inletref::Reference(orc::inletref::prepare(c, n, x3, q));
    report4[0] = r->minimumModeEigenvalue;
    report4[1] = r->minimumPivot;
    report4[2] = r->sourceCharge;
    report4[3] = r->depositedCharge;
    *out = r;
  });
}
void orc_inlet_reference_free(void* h) { delete static_cast<orc::inletref::Reference*>(h); }
// which: 0..5 a native field component, 6 the potential (vertex lattice),
// 7 the interior charge density; count must match.
int orc_inlet_reference_array(void* h, int which, double* out, std::size_t count) {
  return guarded([&] {
    const auto* r = static_cast<const orc::inletref::Reference*>(h);
    const std::vector<double>& v = which < 6 ? r->field[which] : (which == 6 ? r->phi : r->rho);
    if (count != v.size()) throw std::invalid_argument("Inlet reference array size mismatch");
    std::copy(v.begin(), v.end(), out);
  });
}
// movingWindowFieldValue over component a's x-y nodes at fixed-world k.
int orc_inlet_reference_moving(void* h, unsigned a, std::ptrdiff_t k, std::size_t step, double translationBeta,
                               double* plane, std::size_t count) {
  return guarded([&] {
    const auto* r = static_cast<const orc::inletref::Reference*>(h);
    if (a >= 6 || count != r->nx[a] * r->ny[a]) throw std::invalid_argument("Inlet reference plane size mismatch");
    for (std::size_t j = 0; j < r->ny[a]; ++j)
      for (std::size_t i = 0; i < r->nx[a]; ++i)
        plane[j * r->nx[a] + i] = orc::inletref::movingValue(*r, a, i, j, k, step, translationBeta);
  });
}

// EBGaussFieldInitializer::initialize restated (orc_initial_field.hpp);
// the same arguments and report layout as ref_gauss_initialize.
int orc_gauss_initialize(void* h, const double* origin, std::size_t n, const double* x3, const double* u3,
                         const double* q, const double* m, int model, double maxBetaT, const std::size_t* guard,
                         double tolerance, std::size_t maxIterations, int includeCpml, double* report18) {
  return guarded([&] {
    orc::InitialParticles P;
    P.n = n;
    P.x = x3;
    P.u = u3;
    P.q = q;
    P.m = m;
    const orc::GaussReport r = orc::initializeGauss(*F(h), origin, P, static_cast<orc::InitialModel>(model),
                                                    maxBetaT, guard, tolerance, maxIterations, includeCpml != 0);
    const double head[8] = {static_cast<double>(r.iterations), r.relativeResidual, r.maximumGaussResidual,
                            r.totalCharge, r.meanBetaZ, r.meanBetaTransverse, r.initialElectricEnergy,
                            r.initialMagneticEnergy};
    for (int a = 0; a < 8; ++a) report18[a] = head[a];
    for (int a = 0; a < 3; ++a) {
      report18[8 + a] = r.rmsPositionCells[a];
      report18[11 + a] = r.centrePaddingRms[a];
      report18[14 + a] = static_cast<double>(r.staticFieldGuardCells[a]);
    }
    report18[17] = static_cast<double>(r.transversePoissonPaddingCells);
  });
}

// One laboratory field plane (lab_detectors.cpp:2426-2487, one rank).
int orc_lab_plane_new(double zLab, double rhythm, void** out) {
  return guarded([&] {
    // config.cpp:1713-1717
    if (!(rhythm > 0.0) || !std::isfinite(rhythm)) throw std::invalid_argument("field detector rhythm must be positive");
    auto* p = new orc::LabPlaneState;
    p->z = zLab;
    p->rhythm = rhythm;
    *out = p;
  });
}
void orc_lab_plane_free(void* h) { delete static_cast<orc::LabPlaneState*>(h); }
// frame4 = (gamma, lab time origin, lab z origin, box z origin); records:
// nx * ny * 6 doubles; *taken = 1 when this call sampled.
int orc_lab_plane_sample(void* plane, void* fields, const double* frame4, double timeBox, const double* origin,
                         double* records, double* timeLab, int* taken) {
  return guarded([&] {
    const orc::LabFrame f = orc::LabFrame::fromGamma(frame4[0], frame4[1], frame4[2], frame4[3]);
    *taken = static_cast<orc::LabPlaneState*>(plane)->sample(*F(fields), f, timeBox, origin, records, *timeLab) ? 1 : 0;
  });
}

int orc_cowan(double dx, double dy, double dz, double* out12) {
  return guarded([&] {
    const orc::Cowan c = orc::Cowan::forZDispersion(dx, dy, dz);
    for (int a = 0; a < 3; ++a) { out12[a] = c.alpha[a]; out12[3 + a] = c.beta[a]; out12[9 + a] = c.ratio[a]; }
    out12[6] = c.deltaXY; out12[7] = c.deltaYZ; out12[8] = c.deltaZX;
  });
}

int orc_fields_new(std::size_t nx, std::size_t ny, std::size_t nz, double dx, double dy, double dz, double dt,
                   int solver, void** out) {
  return guarded([&] {
    orc::Grid g;
    g.nx = nx; g.ny = ny; g.nz = nz; g.dx = dx; g.dy = dy; g.dz = dz; g.dt = dt; g.solver = solver;
    *out = new orc::Fields(g);
  });
}

void orc_fields_free(void* h) { delete F(h); }

double* orc_fields_component(void* h, int a, std::size_t* count) {
  *count = F(h)->f[a].size();
  return F(h)->f[a].data();
}

int orc_fields_install_cpml(void* h, const std::size_t* cells, double order, double R, double kappa, double alpha,
                            std::size_t globalNz, std::size_t zoff, int lowerZ, int upperZ) {
  return guarded([&] { F(h)->installCpml(params(cells, order, R, kappa, alpha), globalNz, zoff, lowerZ, upperZ); });
}

int orc_fields_install_window(void* h, const std::size_t* cells, double order, double R, double kappa, double alpha) {
  return guarded([&] { F(h)->installWindow(params(cells, order, R, kappa, alpha)); });
}

double* orc_fields_history(void* h, int index, std::size_t* count) {
  std::vector<double>& v = F(h)->history(index);
  *count = v.size();
  return v.data();
}

int orc_fields_advance_magnetic(void* h, const double* const* halo6) {
  return guarded([&] {
    orc::Halo hv;
    if (halo6)
      for (int a = 0; a < 3; ++a) { hv.lower[a] = halo6[a]; hv.upper[a] = halo6[3 + a]; }
    F(h)->advanceMagnetic(hv);
  });
}

int orc_fields_advance_electric(void* h) { return guarded([&] { F(h)->advanceElectric(); }); }
int orc_fields_clear_current(void* h) { return guarded([&] { F(h)->clearCurrent(); }); }

// Window protocol: begin (stages the lower halo), magnetic, [preview], electric, shift.
int orc_window_begin(void* h, std::size_t step, std::ptrdiff_t offset, const double* const* lower3) {
  return guarded([&] { F(h)->beginWindowStep(step, offset, lower3); });
}
// Window with a charge-matched inlet (setNewLayerCompletion with a
// centered normal reference).
int orc_window_set_charge_match(void* h, int enable) {
  return guarded([&] {
    if (!F(h)->window) throw std::logic_error("No moving window is installed");
    F(h)->window->centered = enable != 0;
  });
}
int orc_window_begin_centered(void* h, std::size_t step, std::ptrdiff_t offset, const double* const* lower3,
                              const double* ezAtOffset) {
  return guarded([&] { F(h)->beginWindowStep(step, offset, lower3, ezAtOffset); });
}
int orc_window_advance_magnetic(void* h) {
  return guarded([&] {
    orc::Window& w = *F(h)->window;
    orc::Halo hv;
    for (int a = 0; a < 3; ++a) { hv.lower[a] = w.lower[a].data(); hv.upper[a] = w.upper[a].data(); }
    F(h)->advanceMagnetic(hv);
  });
}
// The inlet overwrite uses reference(a,i,j,offset,step+1) for Ex, Ey.
int orc_window_set_inlet_at(void* h, std::ptrdiff_t offset, std::size_t step1, const double* ex, const double* ey) {
  return guarded([&] {
    const orc::Window& w = *F(h)->window;
    if (offset != w.offset || step1 != w.step + 1) throw std::logic_error("Inlet planes staged for another step");
    F(h)->setWindowInlet(ex, ey);
  });
}
int orc_window_shift(void* h, std::size_t step, std::ptrdiff_t offset, const double* const* planes6) {
  return guarded([&] { F(h)->shiftWindow(step, offset, planes6); });
}
int orc_window_preview(void* h, int enabled) {
  return guarded([&] { F(h)->window->electricPreview = enabled != 0; });
}
int orc_window_undo_preview(void* h) { return guarded([&] { F(h)->undoPreview(); }); }
int orc_window_commit(void* h) {
  return guarded([&] {
    orc::Window& w = *F(h)->window;
    if (w.electricPreview || w.previewPresent) throw std::logic_error("Electric preview not restored before commit");
    F(h)->afterElectricBoundary();
  });
}

int orc_window_enable_ledger(void* h, int enable) {
  return guarded([&] {
    if (!F(h)->window) throw std::logic_error("No moving window is installed");
    if (F(h)->window->established)
      throw std::logic_error("Moving-window energy ledger must be configured before the first step");
    F(h)->window->ledger.enabled = enable != 0;
  });
}
int orc_window_finish_ledger_step(void* h) { return guarded([&] { F(h)->finishEnergyLedgerStep(); }); }
// out: inlet overwrite, inserted, cropped, remap, outward[6], overwrite
// events, flux steps, remap steps.
int orc_window_ledger(void* h, double* out) {
  return guarded([&] {
    orc::Window& w = *F(h)->window;
    if (w.ledger.insertPending && w.phase == 0) F(h)->ledgerInserted();
    const auto& L = w.ledger;
    out[0] = static_cast<double>(L.inletOverwriteEnergyJ);
    out[1] = static_cast<double>(L.insertedEnergyJ);
    out[2] = static_cast<double>(L.croppedEnergyJ);
    out[3] = static_cast<double>(L.remapEnergyJ);
    for (int f = 0; f < 6; ++f) out[4 + f] = static_cast<double>(L.outwardEnergyFace[f]);
    out[10] = static_cast<double>(L.overwriteEvents);
    out[11] = static_cast<double>(L.fluxSteps);
    out[12] = static_cast<double>(L.remapSteps);
  });
}

double orc_fields_energy_density(void* h, std::size_t i, std::size_t j, std::size_t k) {
  return F(h)->energyDensityCell(i, j, k);
}
double orc_fields_poynting(void* h, std::size_t i, std::size_t j, std::size_t k, unsigned axis) {
  return F(h)->poyntingCell(i, j, k, axis);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Particle side.
namespace {
struct CouplingHandle {
  orc::Coupling c;
  std::vector<orc::Particle> particles;
  std::vector<orc::Carrier> carriers;
  orc::BoundaryCounters counters;
  orc::RefCurrent ref;
  orc::HeldRelease held;
  // pending admission queue (rank 0 of the reference) and the step of the
  // last admission
  std::unique_ptr<orc::PendingQueue> pending;
  std::uint64_t pendingStep = 0;
  std::size_t cpmlCells[3] = {0, 0, 0};
  // Particle-boundary geometry: the grid's CPML grading with the particle
  // cells above.
  orc::CpmlParams boundaryParams() const {
    orc::CpmlParams p;
    if (c.fields->cpml) p = c.fields->cpml->p;
    for (int a = 0; a < 3; ++a) p.cells[a] = cpmlCells[a];
    return p;
  }
};
CouplingHandle* CH(void* h) { return static_cast<CouplingHandle*>(h); }
orc::V3 vec(const double* p) { return orc::V3{{p[0], p[1], p[2]}}; }
}  // namespace

extern "C" {

int orc_coupling_new(void* fields, int shape, int smoothX, void** out) {
  return guarded([&] {
    std::unique_ptr<CouplingHandle> h(new CouplingHandle);
    h->c.fields = F(fields);
    h->c.shape = shape;
    h->c.smoothX = smoothX != 0;
    if (smoothX && shape != orc::kBinomial3)
      throw std::invalid_argument("Transverse three-point source requires longitudinal linear-binomial3");
    if (F(fields)->cpml)
      for (int a = 0; a < 3; ++a) h->cpmlCells[a] = F(fields)->cpml->p.cells[a];
    *out = h.release();
  });
}
void orc_coupling_free(void* h) { delete CH(h); }

int orc_coupling_set_frame(void* h, double gamma, double labT0, double labZ0, double boxZ0) {
  return guarded([&] { CH(h)->c.frame.setOriginsFromGamma(gamma, orc::kC, labT0, labZ0, boxZ0); });
}

// type: 0 planar undulator, 1 uniform dipole, 2 quadrupole.
int orc_coupling_add_magnet(void* h, int type, const double* center, double length, double period, double peak,
                            double angle, int gaussianFringe, double cutoff, double gradient, double endRamp,
                            double phase) {
  return guarded([&] {
    orc::Magnet m;
    m.type = type;
    m.center = vec(center);
    m.length = length; m.period = period; m.peak = peak; m.angle = angle;
    m.gaussianFringe = gaussianFringe != 0; m.fringeCutoff = cutoff; m.gradient = gradient;
    m.endRampPeriods = endRamp; m.phase = phase;
    m.prepare();
    CH(h)->c.sources.magnets.push_back(m);
  });
}

int orc_coupling_set_particles(void* h, std::size_t n, const double* x3, const double* u3, const double* q,
                               const double* m, const double* w, const std::uint64_t* id) {
  return guarded([&] {
    auto& P = CH(h)->particles;
    CH(h)->carriers.clear();
    P.assign(n, orc::Particle());
    for (std::size_t i = 0; i < n; ++i) {
      P[i].position = vec(x3 + 3 * i);
      P[i].u = vec(u3 + 3 * i);
      P[i].charge = q[i]; P[i].mass = m[i]; P[i].weight = w[i]; P[i].id = id[i];
    }
  });
}

std::size_t orc_coupling_count(void* h) { return CH(h)->particles.size(); }

int orc_coupling_get_particles(void* h, double* x3, double* u3) {
  return guarded([&] {
    const auto& P = CH(h)->particles;
    for (std::size_t i = 0; i < P.size(); ++i)
      for (int a = 0; a < 3; ++a) { x3[3 * i + a] = P[i].position[a]; u3[3 * i + a] = P[i].u[a]; }
  });
}

// One conservative-source step after beginStep/advanceMagnetic/inlet staging.
// exitLabZ = +inf disables the lab exit plane.
int orc_coupling_step(void* h, const double* origin, double time, unsigned substeps, double exitLabZ,
                      unsigned* iterations, double* force, double* work) {
  return guarded([&] {
    CouplingHandle& H = *CH(h);
    orc::Fields& f = *H.c.fields;
    orc::ParticleBoundary pb;
    pb.set(f.g, vec(origin), H.boundaryParams());
    pb.window = static_cast<bool>(f.window);
    pb.exitLabZ = exitLabZ;
    std::vector<orc::Particle> waiting;
    if (H.pending) waiting = H.pending->near(H.pendingStep, 8);
    const orc::StepResult r = orc::sourceStep(f, H.c, H.particles, H.carriers, H.counters, vec(origin), time,
                                              substeps, pb, &H.ref, H.held, H.pending ? &waiting : nullptr);
    *iterations = r.iterations; *force = r.force; *work = r.work;
  });
}

// Held (frozen) state of the particle list in set_particles order, the
// activation plane (lab z) of config reference.activation_plane_z.
int orc_coupling_set_held(void* h, const int* frozen, const double* beta) {
  return guarded([&] {
    auto& P = CH(h)->particles;
    for (std::size_t i = 0; i < P.size(); ++i) {
      P[i].frozen = frozen[i] != 0;
      P[i].betaHeld = beta[i];
    }
  });
}
int orc_coupling_get_frozen(void* h, int* frozen) {
  return guarded([&] {
    const auto& P = CH(h)->particles;
    for (std::size_t i = 0; i < P.size(); ++i) frozen[i] = P[i].frozen ? 1 : 0;
  });
}
int orc_coupling_set_activation(void* h, double labZ) {
  return guarded([&] { CH(h)->held.activationLabZ = labZ; });
}
int orc_coupling_set_hold_transverse(void* h, int hold) {
  return guarded([&] { CH(h)->held.holdTransverse = hold != 0; });
}

int orc_coupling_get_ids(void* h, std::uint64_t* id) {
  return guarded([&] {
    const auto& P = CH(h)->particles;
    for (std::size_t i = 0; i < P.size(); ++i) id[i] = P[i].id;
  });
}

std::size_t orc_coupling_carrier_count(void* h) { return CH(h)->carriers.size(); }

// Carriers in list order: position, velocity, charge, current weight and the
// id of the particle each came from.
int orc_coupling_get_carriers(void* h, double* x3, double* v3, double* q, double* w, std::uint64_t* id) {
  return guarded([&] {
    const auto& C = CH(h)->carriers;
    for (std::size_t i = 0; i < C.size(); ++i) {
      for (int a = 0; a < 3; ++a) {
        x3[3 * i + a] = C[i].position[a];
        v3[3 * i + a] = C[i].velocity[a];
      }
      q[i] = C[i].charge;
      w[i] = C[i].weight;
      id[i] = C[i].id;
    }
  });
}

// Cumulative boundary counters: counts[20] and charges[20] hold the CPML
// entries per face (0-5), outer exits of particles (6-11) and of carriers
// (12-17), moving-window exits (18) and cropped carriers (19).
int orc_coupling_counters(void* h, std::uint64_t* counts, double* charges) {
  return guarded([&] {
    const auto& k = CH(h)->counters;
    for (int f = 0; f < 6; ++f) {
      counts[f] = k.cpml[f]; charges[f] = k.cpmlQ[f];
      counts[6 + f] = k.direct[f]; charges[6 + f] = k.directQ[f];
      counts[12 + f] = k.carrier[f]; charges[12 + f] = k.carrierQ[f];
    }
    counts[18] = k.exits; charges[18] = k.exitQ;
    counts[19] = k.cropped; charges[19] = k.croppedQ;
  });
}

// The inlet completion of a shift to new_origin (the reference runs it
// inside shift, after the new layer is inserted) over the coupling's
// particles.
int orc_coupling_complete_inlet(void* h, const double* newOrigin) {
  return guarded([&] {
    CouplingHandle& H = *CH(h);
    orc::Fields& f = *H.c.fields;
    if (!f.window || !f.window->centered) throw std::logic_error("The window has no charge-matched inlet");
    std::vector<orc::Particle> sources = H.particles;
    // with the queue: its particles near the inlet at the new step
    // (simulation_moving_window.cpp:317, wide shapes)
    if (H.pending)
      for (const orc::Particle& p : H.pending->near(f.window->step, 8)) sources.push_back(p);
    orc::chargeMatchedInlet(f, sources, vec(newOrigin), H.c.shape, H.c.smoothX, f.cpml->p.cells[0],
                            f.cpml->p.cells[1]);
    if (f.window->ledger.insertPending) f.ledgerInserted();
  });
}

// Energy-ledger particle sums at box time `time` (simulation.cpp:3591-3625).
int orc_coupling_particle_energy(void* h, double time, double* out) {
  return guarded([&] {
    long double sums[9];
    orc::particleEnergySums(CH(h)->particles, CH(h)->c.frame, time, sums);
    for (int c = 0; c < 9; ++c) out[c] = static_cast<double>(sums[c]);
  });
}

// PendingParticleQueue over the coupling's particles (set_particles and
// set_held first): the queue of initial step `initialStep` extracts its
// particles; admit(step) appends the due ones before the step.
int orc_coupling_set_pending(void* h, double referenceOriginZ, std::int64_t initialOffset, double dz,
                             double heldDistancePerBeta, std::uint64_t maximumStep, std::uint64_t initialStep) {
  return guarded([&] {
    CouplingHandle& H = *CH(h);
    orc::PendingGeometry g;
    g.referenceOriginZ = referenceOriginZ;
    g.initialOffset = initialOffset;
    g.dz = dz;
    g.heldDistancePerBeta = heldDistancePerBeta;
    g.maximumStep = maximumStep;
    H.pending.reset(new orc::PendingQueue(g, initialStep));
    H.pending->extract(H.particles);
    H.pendingStep = initialStep;
  });
}
int orc_coupling_admit(void* h, std::uint64_t step, std::size_t* admitted) {
  return guarded([&] {
    CouplingHandle& H = *CH(h);
    if (!H.pending) throw std::logic_error("No pending queue");
    *admitted = H.pending->admit(step, H.particles);
    H.pendingStep = step;
  });
}
// waiting, next due step, admitted
int orc_coupling_pending_state(void* h, std::uint64_t* out3) {
  return guarded([&] {
    CouplingHandle& H = *CH(h);
    if (!H.pending) throw std::logic_error("No pending queue");
    out3[0] = H.pending->size();
    out3[1] = H.pending->nextDueStep();
    out3[2] = H.pending->admitted();
  });
}
// visitNear(step, lookAhead): z and ids of up to `capacity` waiting particles
int orc_coupling_pending_near(void* h, std::uint64_t step, std::uint64_t lookAhead, std::size_t capacity, double* z,
                              std::uint64_t* id, std::size_t* n) {
  return guarded([&] {
    CouplingHandle& H = *CH(h);
    if (!H.pending) throw std::logic_error("No pending queue");
    const std::vector<orc::Particle> near = H.pending->near(step, lookAhead);
    *n = near.size();
    for (std::size_t i = 0; i < near.size() && i < capacity; ++i) {
      z[i] = near[i].position[2];
      id[i] = near[i].id;
    }
  });
}
int orc_repeat_held_addition(double initial, double increment, std::uint64_t count, double* out) {
  return guarded([&] { *out = orc::repeatHeldAddition(initial, increment, count); });
}

// Reference surface current: table of rows x points (the point order of
// TransverseReferenceCurrent's points() for guards gx, gy), translation beta.
int orc_coupling_set_reference_current(void* h, const double* table, std::size_t rows, std::size_t gx,
                                       std::size_t gy, double beta) {
  return guarded([&] {
    CouplingHandle& H = *CH(h);
    H.ref.setPoints(H.c.fields->g, gx, gy);
    H.ref.rows = rows;
    H.ref.table.assign(table, table + rows * H.ref.points.size());
    H.ref.beta = beta;
    H.ref.enabled = true;
  });
}
std::size_t orc_reference_current_points(void* h, std::size_t gx, std::size_t gy, std::size_t* ij) {
  orc::RefCurrent r;
  try {
    r.setPoints(CH(h)->c.fields->g, gx, gy);
  } catch (...) {
    return 0;
  }
  if (ij)
    for (std::size_t p = 0; p < r.points.size(); ++p) { ij[2 * p] = r.points[p].i; ij[2 * p + 1] = r.points[p].j; }
  return r.points.size();
}

// The particle part of finishMovingWindowStep after a shift to new_origin.
int orc_coupling_finish_window(void* h, const double* newOrigin) {
  return guarded([&] {
    CouplingHandle& H = *CH(h);
    orc::finishWindowStep(H.c.fields->g, vec(newOrigin), H.particles, H.carriers, H.counters);
  });
}

// Unit-level boundary geometry of one grid origin: out6 = {hit, fraction,
// face, position[3]} of ParticleCPMLRegion::firstEntry (kind 0),
// ParticleOpenBoundary::firstExit (kind 1); kind 2 returns the damping
// factor currentDampingFactor(a, b, duration) in out6[0].
int orc_boundary_query(void* h, const double* origin, int kind, const double* a, const double* b, double duration,
                       double* out6) {
  return guarded([&] {
    CouplingHandle& H = *CH(h);
    orc::ParticleBoundary pb;
    pb.set(H.c.fields->g, vec(origin), H.boundaryParams());
    orc::Hit hit;
    bool present = false;
    if (kind == 2) {
      out6[0] = pb.damping(vec(a), vec(b), duration, H.c.fields->g);
      return;
    }
    present = kind == 0 ? pb.firstEntry(vec(a), vec(b), hit) : pb.firstExit(vec(a), vec(b), hit);
    out6[0] = present ? 1.0 : 0.0;
    out6[1] = hit.fraction;
    out6[2] = hit.face;
    for (int k = 0; k < 3; ++k) out6[3 + k] = hit.position[k];
  });
}

// Unit-level entry points: the visit list of longitudinal_shape::weights.
int orc_shape_weights(int shape, int smoothX, const double* geometry7, const double* origin, const double* begin,
                      const double* end, int* axis, long long* ijk, double* w, std::size_t capacity,
                      std::size_t* count, unsigned* segments) {
  return guarded([&] {
    orc::Grid g;
    g.nx = (std::size_t)geometry7[0]; g.ny = (std::size_t)geometry7[1]; g.nz = (std::size_t)geometry7[2];
    g.dx = geometry7[3]; g.dy = geometry7[4]; g.dz = geometry7[5]; g.dt = geometry7[6];
    std::size_t n = 0;
    *segments = orc::weights(g, vec(origin), vec(begin), vec(end), [&](unsigned a, long long i, long long j, long long k, double v) {
      if (n >= capacity) throw std::length_error("weights capacity");
      axis[n] = (int)a; ijk[3 * n] = i; ijk[3 * n + 1] = j; ijk[3 * n + 2] = k; w[n] = v; ++n;
    }, shape, smoothX != 0);
    *count = n;
  });
}

int orc_momentum(double* u, const double* e, const double* b, double q, double m, double h, double g0,
                 unsigned* iterations) {
  return guarded([&] {
    orc::V3 uu = vec(u);
    *iterations = orc::momentum(uu, vec(e), vec(b), q, m, h, g0);
    for (int a = 0; a < 3; ++a) u[a] = uu[a];
  });
}

// Copy the current E into the coupling's old-E and set the origin (begin).
int orc_coupling_begin_fields(void* h, const double* origin) {
  return guarded([&] { CH(h)->c.begin(*CH(h)->c.fields, vec(origin), CH(h)->particles); });
}
int orc_coupling_gather(void* h, const double* a, const double* b, double* e, double* bm) {
  return guarded([&] {
    const orc::V3 ee = CH(h)->c.electric(vec(a), vec(b));
    const orc::V3 bb = CH(h)->c.magnetic(vec(a), vec(b));
    for (int k = 0; k < 3; ++k) { e[k] = ee[k]; bm[k] = bb[k]; }
  });
}
int orc_coupling_set_cpml_cells(void* h, const std::size_t* cells) {
  return guarded([&] { for (int a = 0; a < 3; ++a) CH(h)->cpmlCells[a] = cells[a]; });
}
// Shape-level deposit into the grid's J (longitudinal_shape::Current, one rank).
int orc_current_deposit(void* fields, int shape, int smoothX, const double* origin, const double* a, const double* b,
                        double charge) {
  return guarded([&] {
    orc::Coupling c;
    c.fields = F(fields);
    c.shape = shape;
    c.smoothX = smoothX != 0;
    c.origin = vec(origin);
    c.deposit(vec(a), vec(b), charge);
  });
}
int orc_binomial_xz_magnetic(void* fields, const double* origin, const double* point, double* out) {
  return guarded([&] {
    const orc::V3 v = orc::binomialXZMagneticSample(*F(fields), vec(origin), vec(point));
    for (int k = 0; k < 3; ++k) out[k] = v[k];
  });
}
int orc_magnetic_sample_open(void* fields, const double* origin, const double* point, double* out) {
  return guarded([&] {
    const orc::V3 v = orc::magneticSampleOpen(*F(fields), vec(origin), vec(point));
    for (int k = 0; k < 3; ++k) out[k] = v[k];
  });
}
int orc_prescribed_box(void* h, const double* pos, double t, double* e, double* b) {
  return guarded([&] {
    orc::V3 ee = orc::v3(), bb = orc::v3();
    CH(h)->c.sources.addPrescribedBox(vec(pos), t, CH(h)->c.frame, ee, bb);
    for (int k = 0; k < 3; ++k) { e[k] = ee[k]; b[k] = bb[k]; }
  });
}

}  // extern "C"
