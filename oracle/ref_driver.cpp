// Driver around the reference's own hot-path translation units (TEST
// INFRASTRUCTURE ONLY). oracle/Makefile (target ref) compiles this file together with
// the reference sources where they lie under the reference checkout; the
// result goes to oracle/_ref/ and is only used to pin the oracle.
//
// Entry points mirror oracle/orc_capi.cpp with the prefix ref_. The moving
// window's reference callback reads planes staged per (component, k, step),
// so both builds consume identical inlet data.
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include <mpi.h>

#include "compute_threads.h"
#include "eb_cpml.h"
#include "eb_field.h"
#include "moving_window_fields.h"

using namespace aprl;

namespace {
thread_local std::string lastError;

void ensureMpi() {
  static std::once_flag once;
  std::call_once(once, [] {
    int initialized = 0;
    MPI_Initialized(&initialized);
    if (!initialized) {
      int provided = 0;
      MPI_Init_thread(nullptr, nullptr, MPI_THREAD_FUNNELED, &provided);
    }
  });
}

template <class Fn>
int guarded(Fn&& f) {
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

YeeComponent& comp(EBFieldGrid& f, int a) {
  switch (a) {
    case 0: return f.ex(); case 1: return f.ey(); case 2: return f.ez();
    case 3: return f.bx(); case 4: return f.by(); case 5: return f.bz();
    case 6: return f.jx(); case 7: return f.jy(); default: return f.jz();
  }
}

struct RefFields {
  std::unique_ptr<EBFieldGrid> f;
  std::unique_ptr<MovingWindowFields> window;
  std::map<std::tuple<unsigned, std::ptrdiff_t, std::size_t>, std::vector<double>> staged;
  std::size_t planeX[6] = {0, 0, 0, 0, 0, 0};
  EBConvolutionalPML* cpml = nullptr;
  EBCPMLParameters installed;  // grading of the installed CPML (particle damping)
  std::size_t step = 0;

  static Double reference(void* context, unsigned a, std::size_t i, std::size_t j, std::ptrdiff_t k, std::size_t n) {
    RefFields& self = *static_cast<RefFields*>(context);
    auto it = self.staged.find(std::make_tuple(a, k, n));
    if (it == self.staged.end()) throw std::out_of_range("Unstaged moving-window reference plane");
    return it->second[j * self.planeX[a] + i];
  }
  void stage(unsigned a, std::ptrdiff_t k, std::size_t n, const double* plane) {
    const YeeComponent& c = a == 0 ? f->ex() : (a == 1 ? f->ey() : (a == 2 ? f->ez() : (a == 3 ? f->bx() : (a == 4 ? f->by() : f->bz()))));
    staged[std::make_tuple(a, k, n)] = std::vector<double>(plane, plane + c.nx() * c.ny());
  }
};

RefFields* R(void* h) { return static_cast<RefFields*>(h); }

EBCPMLParameters params(const std::size_t cells[3], double order, double R, double kappa, double alpha) {
  EBCPMLParameters p;
  for (int a = 0; a < 3; ++a) p.cells[a] = cells[a];
  p.polynomialOrder = order;
  p.targetReflection = R;
  p.kappaMax = kappa;
  p.alphaFraction = alpha;
  return p;
}
}  // namespace

extern "C" {

const char* ref_last_error() { return lastError.c_str(); }


// configureComputeThreads (compute_threads.cpp:50-91) for the CPU baseline.
int ref_set_threads(int threads) {
  return guarded([&] {
    ensureMpi();
    configureComputeThreads(MPI_THREAD_FUNNELED, threads);
  });
}

int ref_cowan(double dx, double dy, double dz, double* out12) {
  return guarded([&] {
    const EBCowanCoefficients c = EBCowanCoefficients::forZDispersion(dx, dy, dz);
    for (int a = 0; a < 3; ++a) { out12[a] = c.alpha[a]; out12[3 + a] = c.beta[a]; out12[9 + a] = c.ratio[a]; }
    out12[6] = c.deltaXY; out12[7] = c.deltaYZ; out12[8] = c.deltaZX;
  });
}

int ref_fields_new(std::size_t nx, std::size_t ny, std::size_t nz, double dx, double dy, double dz, double dt,
                   int solver, void** out) {
  return guarded([&] {
    ensureMpi();
    std::unique_ptr<RefFields> r(new RefFields);
    r->f.reset(new EBFieldGrid(EBGridGeometry(nx, ny, nz, dx, dy, dz, dt,
                                              solver ? EBMaxwellSolver::CowanZ : EBMaxwellSolver::Yee)));
    for (int a = 0; a < 6; ++a) r->planeX[a] = comp(*r->f, a).nx();
    *out = r.release();
  });
}

void ref_fields_free(void* h) { delete R(h); }

double* ref_fields_component(void* h, int a, std::size_t* count) {
  YeeComponent& c = comp(*R(h)->f, a);
  *count = c.size();
  return c.data();
}

int ref_fields_install_cpml(void* h, const std::size_t* cells, double order, double Rf, double kappa, double alpha,
                            std::size_t globalNz, std::size_t zoff, int lowerZ, int upperZ) {
  return guarded([&] {
    std::unique_ptr<EBConvolutionalPML> p(new EBConvolutionalPML(R(h)->f->geometry(), globalNz, zoff,
                                                                 params(cells, order, Rf, kappa, alpha), lowerZ != 0,
                                                                 upperZ != 0));
    R(h)->cpml = p.get();
    R(h)->installed = params(cells, order, Rf, kappa, alpha);
    R(h)->f->setBoundary(std::move(p));
  });
}

int ref_fields_install_window(void* h, const std::size_t* cells, double order, double Rf, double kappa, double alpha) {
  return guarded([&] {
    RefFields& r = *R(h);
    MovingWindowReference ref;
    ref.context = &r;
    ref.value = &RefFields::reference;
    const EBGridGeometry& g = r.f->geometry();
    r.window.reset(new MovingWindowFields(g, g.nz, 0, MPI_COMM_SELF, params(cells, order, Rf, kappa, alpha), ref));
    r.window->installBoundary(*r.f);
    r.cpml = dynamic_cast<EBConvolutionalPML*>(&r.f->boundary());
    r.installed = params(cells, order, Rf, kappa, alpha);
  });
}

double* ref_fields_history(void* h, int index, std::size_t* count) {
  std::vector<double>& v = R(h)->cpml->checkpointHistory(index);
  *count = v.size();
  return v.data();
}

int ref_fields_advance_magnetic(void* h, const double* const* halo6) {
  return guarded([&] {
    EBElectricHaloView hv;
    if (halo6) {
      hv.lowerEx = halo6[0]; hv.lowerEy = halo6[1]; hv.lowerEz = halo6[2];
      hv.upperEx = halo6[3]; hv.upperEy = halo6[4]; hv.upperEz = halo6[5];
    }
    R(h)->f->advanceMagnetic(hv);
  });
}

int ref_fields_advance_electric(void* h) { return guarded([&] { R(h)->f->advanceElectric(); }); }
int ref_fields_clear_current(void* h) { return guarded([&] { R(h)->f->clearCurrent(); }); }

int ref_window_begin(void* h, std::size_t step, std::ptrdiff_t offset, const double* const* lower3) {
  return guarded([&] {
    RefFields& r = *R(h);
    r.staged.clear();
    for (unsigned a = 0; a < 3; ++a) r.stage(a, offset - 1, step, lower3[a]);
    r.step = step;
    r.window->beginStep(*r.f, step, offset);
  });
}
int ref_window_advance_magnetic(void* h) {
  return guarded([&] {
    RefFields& r = *R(h);
    const EBElectricHaloView hv = r.window->physicalElectricHalos(*r.f, r.step);
    r.f->advanceMagnetic(hv);
  });
}
// The inlet overwrite queries reference(a,i,j,offset,step+1) for Ex, Ey.
int ref_window_set_inlet_at(void* h, std::ptrdiff_t offset, std::size_t step1, const double* ex, const double* ey) {
  return guarded([&] {
    R(h)->stage(0, offset, step1, ex);
    R(h)->stage(1, offset, step1, ey);
  });
}
int ref_window_shift(void* h, std::size_t step, std::ptrdiff_t offset, const double* const* planes6) {
  return guarded([&] {
    RefFields& r = *R(h);
    for (unsigned a = 0; a < 6; ++a) r.stage(a, offset, step, planes6[a]);
    r.window->shift(*r.f, step, offset);
  });
}
int ref_window_preview(void* h, int enabled) {
  return guarded([&] { R(h)->window->previewElectricBoundary(enabled != 0); });
}
int ref_window_undo_preview(void* h) { return guarded([&] { R(h)->window->undoPreviewElectricBoundary(*R(h)->f); }); }
int ref_window_commit(void* h) { return guarded([&] { R(h)->window->commitPreparedElectricBoundary(*R(h)->f); }); }
int ref_window_enable_ledger(void* h, int enable) {
  return guarded([&] { R(h)->window->enableEnergyLedger(enable != 0); });
}
int ref_window_finish_ledger_step(void* h) {
  return guarded([&] { R(h)->window->finishEnergyLedgerStep(*R(h)->f); });
}
int ref_window_ledger(void* h, double* out) {
  return guarded([&] {
    const MovingWindowFieldEnergyLedger& L = R(h)->window->energyLedger();
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

double ref_fields_energy_density(void* h, std::size_t i, std::size_t j, std::size_t k) {
  return R(h)->f->fieldEnergyDensityCell(i, j, k);
}
double ref_fields_poynting(void* h, std::size_t i, std::size_t j, std::size_t k, unsigned axis) {
  return R(h)->f->poyntingComponentCell(i, j, k, axis);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Particle side: the reference's ConservativeSourceCoupling, shapes and
// sources driven by a restatement of simulation_source_coupling.cpp:16-65 and
// the active-particle branch of simulation.cpp:4167-5605.
#include <cstdlib>
#include <unistd.h>

#include "boostframe.h"
#include "eb_bunch.h"
#include "eb_initial_field.h"
#include "compute_threads.h"
#include "eb_deposition.h"
#include "eb_particles.h"
#include "eb_sources.h"
#include "longitudinal_shape.h"
#include "particle_boundary.h"
#include "source_coupling.h"
#include "source_coupling_math.h"
#include "charge_matched_inlet.h"
#include "pending_particle_queue.h"

namespace {
struct RefCoupling {
  RefFields* fields = nullptr;
  std::unique_ptr<ConservativeSourceCoupling> c;
  SIFieldSourceSet sources;
  BoostFrameTransform frame;
  std::vector<RelativisticParticleSI> particles;
  std::vector<ParticleCPMLCarrier> carriers;
  std::vector<std::uint64_t> carrierIds;  // bookkeeping: the particle each carrier came from
  // cumulative counters: CPML entries (0-5), particle outer exits (6-11),
  // carrier outer exits (12-17), window exits (18), cropped carriers (19)
  unsigned long long counts[20] = {};
  double charges[20] = {};
  std::string dir;
  bool quintic = false, binomial3 = false, smoothX = false;
  EBCPMLParameters cpml;
  Double activationLabZ = std::numeric_limits<Double>::infinity();
  bool holdTransverse = true;  // reference.hold_transverse_until_activation
  // rank 0's pending admission queue and the step of its last admission
  std::unique_ptr<PendingParticleQueue> pending;
  std::size_t pendingStep = 0;
  ~RefCoupling() {
    c.reset();
    if (!dir.empty()) {
      std::string cmd = "rm -rf '" + dir + "'";
      if (std::system(cmd.c_str())) {}
    }
  }
};
RefCoupling* RC(void* h) { return static_cast<RefCoupling*>(h); }
FieldVector<Double> fv(const double* p) {
  FieldVector<Double> v(0.0);
  for (int a = 0; a < 3; ++a) v[a] = p[a];
  return v;
}
}  // namespace

extern "C" {

// The laboratory-plane input route through the reference's own
// SIBunchPreprocessor (eb_bunch.cpp) on MPI_COMM_SELF; the frozen-injection
// frame origin and the frozen flags (simulation.cpp:2126-2186, 2228-2245,
// not buildable here) are transcribed around the reference's own
// BoostFrameTransform. Arguments and reports as the orc_bunch_* entry points.
std::vector<RelativisticParticleSI> refParticles(std::size_t n, const double* x3, const double* u3, const double* w) {
  std::vector<RelativisticParticleSI> P(n);
  for (std::size_t p = 0; p < n; ++p) {
    for (int a = 0; a < 3; ++a) {
      P[p].position[a] = x3[3 * p + a];
      P[p].properVelocity[a] = u3 ? u3[3 * p + a] : 0.0;
    }
    P[p].weight = w ? w[p] : 1.0;
  }
  return P;
}
void refStore(const std::vector<RelativisticParticleSI>& P, double* x3, double* u3) {
  for (std::size_t p = 0; p < P.size(); ++p)
    for (int a = 0; a < 3; ++a) {
      x3[3 * p + a] = P[p].position[a];
      if (u3) u3[3 * p + a] = P[p].properVelocity[a];
    }
}
int ref_bunch_project(std::size_t n, double* x3, const double* u3, const double* w, double inputPlane,
                      double reference, double offset, int hold, double* report8) {
  return guarded([&] {
    ensureMpi();
    auto P = refParticles(n, x3, u3, w);
    const SILabPlaneProjectionReport r = SIBunchPreprocessor::projectLabPlaneEventsToSnapshot(
        P, inputPlane, reference, offset, MPI_COMM_SELF, hold != 0);
    refStore(P, x3, nullptr);
    const double v[8] = {r.inputPlaneLab, r.referencePositionLab, r.minimumForwardDistance, r.maximumForwardDistance,
                         r.recommendedMinimumReferencePosition, r.referenceTimeOffsetLab, r.meanLongitudinalBeta,
                         static_cast<double>(r.particles)};
    for (int a = 0; a < 8; ++a) report8[a] = v[a];
  });
}
int ref_bunch_place(std::size_t n, double* x3, int hasEntrance, double entrance, double reference, double margin,
                    double* report7) {
  return guarded([&] {
    ensureMpi();
    auto P = refParticles(n, x3, nullptr, nullptr);
    SIBunchPlacement placement;
    placement.hasFirstInteractionEntrance = hasEntrance != 0;
    placement.firstInteractionEntranceLab = entrance;
    placement.referencePositionLab = reference;
    placement.recommendationMarginLab = margin;
    const SIBunchPlacementReport r = SIBunchPreprocessor::placeRelativeLabSnapshot(P, placement, MPI_COMM_SELF);
    refStore(P, x3, nullptr);
    const double v[7] = {r.relativeTailLab, r.relativeHeadLab, r.referencePositionLab, r.headAfterLab,
                         r.actualHeadDistance, r.recommendedMaximumReferencePosition,
                         static_cast<double>(r.particles)};
    for (int a = 0; a < 7; ++a) report7[a] = v[a];
  });
}
int ref_bunch_frozen_origin(std::size_t n, const double* x3, const double* u3, double gammaBoost, double inputPlane,
                            double activationPlane, double dz, int hold, double* out2) {
  return guarded([&] {
    const auto P = refParticles(n, x3, u3, nullptr);
    Double earliest = std::numeric_limits<Double>::infinity();
    int invalidHeld = 0;
    BoostFrameTransform probe;
    if (hold) probe.setOriginsFromGamma(gammaBoost, SI::c, 0.0, activationPlane, 0.0);
    for (const RelativisticParticleSI& particle : P) {
      const Double gamma = BoostFrameTransform::gammaFromProperVelocity(particle.properVelocity);
      const Double betaZ = particle.properVelocity[2] / gamma;
      const Double snapshotZ = particle.position[2];
      if (hold) {
        FieldVector<Double> uBox(0.0);
        probe.properVelocityLabToBox(particle.properVelocity, uBox);
        const Double betaHeld = probe.heldLongitudinalBetaZ(uBox);
        const Double labAdvance = probe.beta() + betaHeld;
        if (!(labAdvance > 0.0) || !std::isfinite(labAdvance)) {
          invalidHeld = 1;
          continue;
        }
        const Double inputTime = (inputPlane - snapshotZ) / (betaZ * SI::c);
        const Double inputToActivation = inputPlane - activationPlane;
        const Double crossing = inputTime - inputToActivation * (1.0 + probe.beta() * betaHeld) / (SI::c * labAdvance);
        const Double margin = 2.0 * dz / (probe.gamma() * SI::c * labAdvance);
        earliest = std::min(earliest, crossing - margin);
      } else {
        earliest = std::min(earliest, (activationPlane - snapshotZ) / (betaZ * SI::c));
      }
    }
    if (invalidHeld)
      throw std::invalid_argument("Held-transverse pseudo-trajectory must advance in laboratory z for every particle");
    out2[0] = hold ? earliest : earliest - 2.0 * dz / (gammaBoost * SI::c);
    out2[1] = activationPlane;
  });
}
int ref_bunch_boost(std::size_t n, double* x3, double* u3, const double* frame4, double snapshotTime, int hold,
                    double inputPlane, double* report7) {
  return guarded([&] {
    auto P = refParticles(n, x3, u3, nullptr);
    BoostFrameTransform frame;
    frame.setOriginsFromGamma(frame4[0], SI::c, frame4[1], frame4[2], frame4[3]);
    const SIBunchBoostReport r =
        SIBunchPreprocessor::boostLabSnapshotToBoxTimeZero(P, frame, snapshotTime, hold != 0, inputPlane);
    refStore(P, x3, u3);
    const double v[7] = {r.earliestLabEventTime, r.latestLabEventTime, r.maximumAbsoluteDriftTime,
                         r.minimumLabGamma, r.maximumLabGamma, r.maximumRelativeMomentumRoundTripError,
                         r.maximumRelativeGammaRoundTripError};
    for (int a = 0; a < 7; ++a) report7[a] = v[a];
  });
}
int ref_bunch_freeze(std::size_t n, const double* x3, const double* u3, const double* frame4, double activationPlane,
                     int hold, double* beta) {
  return guarded([&] {
    auto P = refParticles(n, x3, u3, nullptr);
    BoostFrameTransform frame;
    frame.setOriginsFromGamma(frame4[0], SI::c, frame4[1], frame4[2], frame4[3]);
    int alreadyPast = 0;
    for (RelativisticParticleSI& particle : P) {
      particle.frozen = true;
      if (hold) particle.frozenBetaZ = frame.heldLongitudinalBetaZ(particle.properVelocity);
      if (frame.labZFromBoxZT(particle.position[2], 0.0) >= activationPlane) alreadyPast = 1;
    }
    if (alreadyPast)
      throw std::runtime_error(
          "Frozen lab-plane injection requires every particle upstream of the input plane at box time zero");
    for (std::size_t p = 0; p < n; ++p) beta[p] = hold ? P[p].frozenBetaZ : 0.0;
  });
}

// The reference's BoostFrameTransform for the detector: out = {box z of the
// lab plane, its lab time, E_lab[3], B_lab[3]}.
// EBGaussFieldInitializer::initialize (eb_initial_field.cpp:829-1095) on
// one rank; report18 = iterations, relative residual, maximum Gauss
// residual, total charge, mean beta_z, mean beta_transverse, E and B
// energies, rms position (cells, 3), centre padding (rms, 3), guard cells
// (3), transverse padding.
int ref_gauss_initialize(void* h, const double* origin, std::size_t n, const double* x3, const double* u3,
                         const double* q, const double* m, int model, double maxBetaT, const std::size_t* guard,
                         double tolerance, std::size_t maxIterations, int includeCpml, double* report18) {
  return guarded([&] {
    ensureMpi();
    EBFieldGrid& f = *R(h)->f;
    std::vector<RelativisticParticleSI> particles(n);
    for (std::size_t p = 0; p < n; ++p) {
      for (int a = 0; a < 3; ++a) {
        particles[p].position[a] = x3[3 * p + a];
        particles[p].properVelocity[a] = u3[3 * p + a];
      }
      particles[p].charge = q[p];
      particles[p].mass = m[p];
      particles[p].weight = 1.0;
      particles[p].id = p + 1;
    }
    FieldVector<Double> o(0.0);
    for (int a = 0; a < 3; ++a) o[a] = origin[a];
    const InitialSelfFieldModel models[3] = {InitialSelfFieldModel::RelativisticPoisson,
                                             InitialSelfFieldModel::ElectrostaticPoisson,
                                             InitialSelfFieldModel::CowanRelativisticPoisson};
    const EBGaussInitializationReport r = EBGaussFieldInitializer::initialize(
        f, f.geometry(), 0, o, particles, models[model], maxBetaT, guard, tolerance, maxIterations, MPI_COMM_WORLD,
        includeCpml != 0, false, nullptr);
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

int ref_boost_lab(double gamma, double labT0, double labZ0, double boxZ0, double zLab, double tBox,
                  const double* eBox, const double* bBox, double* out8) {
  return guarded([&] {
    BoostFrameTransform f;
    f.setOriginsFromGamma(gamma, SI::c, labT0, labZ0, boxZ0);
    out8[0] = f.boxZFromLabZAndBoxT(zLab, tBox);
    out8[1] = f.labTFromLabZBoxT(zLab, tBox);
    FieldVector<Double> e(0.0), b(0.0), el(0.0), bl(0.0);
    for (int a = 0; a < 3; ++a) {
      e[a] = eBox[a];
      b[a] = bBox[a];
    }
    f.fieldsBoxToLab(e, b, el, bl);
    for (int a = 0; a < 3; ++a) {
      out8[2 + a] = el[a];
      out8[5 + a] = bl[a];
    }
  });
}

int ref_coupling_new(void* fields, int shape, int smoothX, void** out) {
  return guarded([&] {
    std::unique_ptr<RefCoupling> h(new RefCoupling);
    h->fields = R(fields);
    h->quintic = shape == 0;
    h->binomial3 = shape == 2;
    h->smoothX = smoothX != 0;
    char tmpl[] = "/tmp/aprl_ref_XXXXXX";
    if (!mkdtemp(tmpl)) throw std::runtime_error("Cannot create a coupling log directory");
    h->dir = tmpl;
    h->c.reset(new ConservativeSourceCoupling(MPI_COMM_SELF, h->dir, h->quintic, shape == 1, h->binomial3, false, false,
                                              h->smoothX));
    *out = h.release();
  });
}
void ref_coupling_free(void* h) { delete RC(h); }

int ref_coupling_set_frame(void* h, double gamma, double labT0, double labZ0, double boxZ0) {
  return guarded([&] { RC(h)->frame.setOriginsFromGamma(gamma, SI::c, labT0, labZ0, boxZ0); });
}

int ref_coupling_add_magnet(void* h, int type, const double* center, double length, double period, double peak,
                            double angle, int gaussianFringe, double cutoff, double gradient, double endRamp,
                            double phase) {
  return guarded([&] {
    SIMagneticElement m;
    m.type = type == 0 ? SIMagnetType::PlanarUndulator : (type == 1 ? SIMagnetType::UniformDipole : SIMagnetType::Quadrupole);
    m.center = fv(center);
    m.length = length; m.period = period; m.peakMagneticField = peak; m.polarizationAngle = angle;
    m.gaussianFringe = gaussianFringe != 0; m.fringeRelativeCutoff = cutoff; m.quadrupoleGradient = gradient;
    m.endRampPeriods = endRamp; m.undulatorPhase = phase;
    RC(h)->sources.addPrescribedLabMagnet(m);
  });
}

int ref_coupling_set_cpml_cells(void* h, const std::size_t* cells) {
  return guarded([&] { for (int a = 0; a < 3; ++a) RC(h)->cpml.cells[a] = cells[a]; });
}

int ref_coupling_set_particles(void* h, std::size_t n, const double* x3, const double* u3, const double* q,
                               const double* m, const double* w, const std::uint64_t* id) {
  return guarded([&] {
    auto& P = RC(h)->particles;
    P.assign(n, RelativisticParticleSI());
    for (std::size_t i = 0; i < n; ++i) {
      P[i].position = fv(x3 + 3 * i);
      P[i].properVelocity = fv(u3 + 3 * i);
      P[i].charge = q[i]; P[i].mass = m[i]; P[i].weight = w[i]; P[i].id = id[i];
      P[i].frozen = false;
    }
    RC(h)->carriers.clear();
    RC(h)->carrierIds.clear();
  });
}
std::size_t ref_coupling_count(void* h) { return RC(h)->particles.size(); }
int ref_coupling_get_particles(void* h, double* x3, double* u3) {
  return guarded([&] {
    const auto& P = RC(h)->particles;
    for (std::size_t i = 0; i < P.size(); ++i)
      for (int a = 0; a < 3; ++a) { x3[3 * i + a] = P[i].position[a]; u3[3 * i + a] = P[i].properVelocity[a]; }
  });
}

int ref_coupling_step(void* h, const double* origin, double time, unsigned substeps, double exitLabZ,
                      unsigned* iterations, double* force, double* work) {
  return guarded([&] {
    RefCoupling& H = *RC(h);
    RefFields& rf = *H.fields;
    EBFieldGrid& f = *rf.f;
    const EBGridGeometry& g = f.geometry();
    const FieldVector<Double> o = fv(origin);
    EBCPMLParameters boundaryParams = rf.installed;
    for (int a = 0; a < 3; ++a) boundaryParams.cells[a] = H.cpml.cells[a];
    ParticleCPMLRegion cpmlRegion(g, o, boundaryParams);
    ParticleOpenBoundary outer(g, o, g, o, 0);
    const bool window = static_cast<bool>(rf.window);
    auto pushDepositAll = [&]() {
      ChargeConservingCurrentDepositor depositor(f, o, H.c->longitudinalCurrent());
      // pushDepositAndMigrate's pending support (simulation.cpp:4170-4172)
      if ((H.quintic || H.binomial3) && H.pending)
        H.pending->visitNear(H.pendingStep, 8, [&](const RelativisticParticleSI& p) {
          auto end = p.position;
          end[2] += SI::c * g.dt * p.frozenBetaZ;
          H.c->depositFrozenSourceSupport(p.position, end, p.charge, false);
        });
      std::vector<RelativisticParticleSI> prepared = H.particles;
      // prepareParticle (simulation.cpp:4919-5035), hold-transverse release
      std::vector<char> hasActive(prepared.size(), 0);
      std::vector<Double> activeTime(prepared.size(), time);
      std::vector<FieldVector<Double>> activePosition(prepared.size(), FieldVector<Double>(0.0));
      auto prepare = [&](std::size_t n) {
        RelativisticParticleSI& particle = prepared[n];
        const FieldVector<Double> start(particle.position);
        if (window && start[2] < o[2] && !particle.frozen)
          throw std::runtime_error("An active particle is upstream of the moving window; the inlet guard did not contain its release event");
        hasActive[n] = !particle.frozen;
        activePosition[n] = particle.position;
        if (particle.frozen) {
          const Double betaHeld = H.holdTransverse ? particle.frozenBetaZ : 0.0;
          const Double startLabZ = H.frame.labZFromBoxZT(start[2], time);
          if (H.holdTransverse)
            particle.position[2] += SI::c * g.dt * betaHeld;
          else
            RelativisticBorisPusher::pushPosition(particle.position, particle.properVelocity, g.dt);
          const Double endLabZ = H.frame.labZFromBoxZT(particle.position[2], time + g.dt);
          if (endLabZ >= H.activationLabZ) {
            if (window && start[2] < o[2])
              throw std::runtime_error("A pending upstream held particle would release before a complete in-window step; check the moving inlet guard");
            const Double crossingFraction = startLabZ >= H.activationLabZ ? 0.0 :
              std::max(0.0, std::min(1.0, (H.activationLabZ - startLabZ) / (endLabZ - startLabZ)));
            particle.position = start;
            if (H.holdTransverse)
              particle.position[2] += SI::c * crossingFraction * g.dt * betaHeld;
            else
              RelativisticBorisPusher::pushPosition(particle.position, particle.properVelocity,
                                                    crossingFraction * g.dt);
            particle.frozen = false;
            hasActive[n] = 1;
            activePosition[n] = particle.position;
            activeTime[n] = time + crossingFraction * g.dt;
            const Double remaining = (1.0 - crossingFraction) * g.dt;
            if (remaining > 0.0)
              H.c->push(particle, H.sources, H.frame, activeTime[n], remaining, substeps, nullptr, nullptr, nullptr);
          }
        } else {
          H.c->push(particle, H.sources, H.frame, time, g.dt, substeps, nullptr, nullptr, nullptr);
        }
      };
      // Preparation in batches of 2048 with OpenMP, as simulation.cpp:5041-5080.
      const std::ptrdiff_t total = static_cast<std::ptrdiff_t>(prepared.size());
      for (std::ptrdiff_t batch = 0; batch < total; batch += 2048) {
        const std::ptrdiff_t count = std::min<std::ptrdiff_t>(2048, total - batch);
        if (computeThreadCount() == 1 || count < 32) {
          for (std::ptrdiff_t k = 0; k < count; ++k) prepare(static_cast<std::size_t>(batch + k));
        } else {
          std::exception_ptr error;
#pragma omp parallel for schedule(static) num_threads(computeThreadCount())
          for (std::ptrdiff_t k = 0; k < count; ++k) {
            try {
              prepare(static_cast<std::size_t>(batch + k));
            } catch (...) {
#pragma omp critical(aprl_ref_prepare_error)
              if (!error) error = std::current_exception();
            }
          }
          if (error) std::rethrow_exception(error);
        }
      }
      // processCarrierSegment (simulation.cpp:4310-4407), one rank.
      const Double cutoff = 64.0 * std::numeric_limits<Double>::epsilon();
      std::vector<ParticleCPMLCarrier> keptCarriers;
      std::vector<std::uint64_t> keptIds;
      auto carrierSegment = [&](ParticleCPMLCarrier carrier, std::uint64_t id, const FieldVector<Double>& segStart,
                                Double startFraction) {
        if (carrier.currentWeight <= cutoff) return;
        ParticleBoundaryHit outerHit;
        const bool exits = outer.firstExit(segStart, carrier.position, outerHit);
        Double localEnd = 1.0;
        FieldVector<Double> segEnd(carrier.position);
        if (exits) {
          localEnd = outerHit.fraction;
          segEnd = outerHit.position;
        }
        const Double globalEnd = startFraction + localEnd * (1.0 - startFraction);
        const Double duration = (globalEnd - startFraction) * g.dt;
        FieldVector<Double> mid(0.0);
        for (unsigned a = 0; a < 3; ++a) mid[a] = 0.5 * (segStart[a] + segEnd[a]);
        const Double midDamping = cpmlRegion.currentDampingFactor(segStart, mid, 0.5 * duration);
        const Double endDamping = cpmlRegion.currentDampingFactor(segStart, segEnd, duration);
        const Double endWeight = carrier.currentWeight * endDamping;
        depositor.depositSegment(segStart, segEnd, carrier.charge * (carrier.currentWeight * midDamping));
        if (exits) {
          outer.depositOutgoingFlux(outerHit, carrier.charge * endWeight);
          const std::size_t face = static_cast<std::size_t>(outerHit.face);
          ++H.counts[12 + face];
          H.charges[12 + face] += carrier.charge * endWeight;
          return;
        }
        if (endWeight <= cutoff) return;
        carrier.currentWeight = endWeight;
        carrier.position = segEnd;
        keptCarriers.push_back(carrier);
        keptIds.push_back(id);
      };
      for (std::size_t n = 0; n < H.carriers.size(); ++n) {
        ParticleCPMLCarrier carrier = H.carriers[n];
        const FieldVector<Double> start(carrier.position);
        carrier.position.pmv(g.dt, carrier.velocity);
        carrierSegment(carrier, H.carrierIds[n], start, 0.0);
      }
      // processActiveSegment (simulation.cpp:4555-4917) for active particles.
      const Double tolerance = 64.0 * std::numeric_limits<Double>::epsilon();
      std::vector<RelativisticParticleSI> retained;
      for (std::size_t n = 0; n < H.particles.size(); ++n) {
        FieldVector<Double> start(H.particles[n].position);
        Double startFraction = 0.0;
        const RelativisticParticleSI& p = prepared[n];
        // moving-window upstream admission (simulation.cpp:5121-5155)
        if (window && start[2] < o[2]) {
          if (!H.particles[n].frozen || !p.frozen || hasActive[n])
            throw std::runtime_error("Moving-window upstream admission requires an unreleased held particle on rank zero");
          if (p.position[2] < o[2]) {
            H.c->depositFrozenSourceSupport(start, p.position, p.charge, false);
            retained.push_back(p);
            continue;
          }
          const Double fraction = (o[2] - start[2]) / (p.position[2] - start[2]);
          if (!(fraction >= 0.0 && fraction <= 1.0) || !std::isfinite(fraction))
            throw std::runtime_error("Invalid moving-window admission fraction");
          for (unsigned axis = 0; axis < 3; ++axis) start[axis] += fraction * (p.position[axis] - start[axis]);
          start[2] = o[2];
          startFraction = fraction;
          H.c->depositFrozenSourceSupport(H.particles[n].position, start, p.charge, false);
        }
        const bool kink = H.particles[n].frozen && hasActive[n] && activeTime[n] > time && activeTime[n] < time + g.dt;
        if (kink) {
          // processActiveSegment's frozen leg (simulation.cpp:4676-4746)
          ParticleBoundaryHit h1, h2;
          if (cpmlRegion.firstEntry(start, activePosition[n], h1) || outer.firstExit(start, activePosition[n], h2))
            throw std::runtime_error("Held-transverse activation crossed a particle boundary during the frozen part of a field step; enable mesh.ballistic_outside_transverse for boundary injection");
          depositor.depositSegment(start, activePosition[n], p.charge);
          start = activePosition[n];
          startFraction = (activeTime[n] - time) / g.dt;
        }
        ParticleBoundaryHit cpmlHit, outerHit;
        // firstMovingWindowEntry (simulation_moving_window.cpp:380-395)
        bool entersCpml = cpmlRegion.firstEntry(start, p.position, cpmlHit);
        if (window && std::isfinite(exitLabZ)) {
          const double a = H.frame.labZFromBoxZT(start[2], time + startFraction * g.dt) - exitLabZ;
          const double b = H.frame.labZFromBoxZT(p.position[2], time + g.dt) - exitLabZ;
          if (a >= 0.) throw std::runtime_error("Active particle started beyond trusted moving-window exit");
          if (b >= 0. && b > a) {
            const double fraction = -a / (b - a);
            if (!entersCpml || fraction < cpmlHit.fraction) {
              for (unsigned d = 0; d < 3; ++d) cpmlHit.position[d] = start[d] + fraction * (p.position[d] - start[d]);
              cpmlHit.fraction = fraction;
              cpmlHit.face = ParticleBoundaryFace::UpperZ;
              entersCpml = true;
            }
          }
        }
        const bool exitsOuter = outer.firstExit(start, p.position, outerHit);
        int action = 0;  // retain / enter CPML / exit outer
        Double localEnd = 1.0;
        FieldVector<Double> segEnd(p.position);
        ParticleBoundaryHit selected;
        if (entersCpml && (!exitsOuter || cpmlHit.fraction <= outerHit.fraction + tolerance)) {
          action = 1;
          localEnd = cpmlHit.fraction;
          segEnd = cpmlHit.position;
          selected = cpmlHit;
        } else if (exitsOuter) {
          action = 2;
          localEnd = outerHit.fraction;
          segEnd = outerHit.position;
          selected = outerHit;
        }
        depositor.depositSegment(start, segEnd, p.charge);
        const std::size_t face = static_cast<std::size_t>(selected.face);
        if (action == 1) {
          if (window && selected.face == ParticleBoundaryFace::UpperZ) {
            ++H.counts[18];
            H.charges[18] += p.charge;
          } else {
            ++H.counts[face];
            H.charges[face] += p.charge;
          }
          ParticleCPMLCarrier carrier;
          carrier.position = p.position;
          const Double gamma = BoostFrameTransform::gammaFromProperVelocity(p.properVelocity);
          carrier.velocity = p.properVelocity;
          carrier.velocity *= SI::c / gamma;
          carrier.charge = p.charge;
          carrier.currentWeight = 1.0;
          carrierSegment(carrier, p.id, selected.position, startFraction + localEnd * (1.0 - startFraction));
        } else if (action == 2) {
          outer.depositOutgoingFlux(selected, p.charge);
          ++H.counts[6 + face];
          H.charges[6 + face] += p.charge;
        } else {
          retained.push_back(p);
        }
      }
      H.particles.swap(retained);
      H.carriers.swap(keptCarriers);
      H.carrierIds.swap(keptIds);
    };
    ConservativeSourceCoupling& coupling = *H.c;
    coupling.begin(f, o, H.particles, H.cpml.cells[0], H.cpml.cells[1]);
    rf.window->previewElectricBoundary(true);
    f.advanceElectric();
    rf.window->previewElectricBoundary(false);
    const auto original = H.particles;
    const auto originalCarriers = H.carriers;
    const auto originalIds = H.carrierIds;
    unsigned long long counts0[20];
    double charges0[20];
    std::copy(H.counts, H.counts + 20, counts0);
    std::copy(H.charges, H.charges + 20, charges0);
    auto restore = [&]() {
      H.particles = original;
      H.carriers = originalCarriers;
      H.carrierIds = originalIds;
      std::copy(counts0, counts0 + 20, H.counts);
      std::copy(charges0, charges0 + 20, H.charges);
    };
    bool converged = false;
    unsigned its = 0;
    Double frc = 0., wrk = 0.;
    for (unsigned n = 0; n < 64; ++n) {
      if (n) coupling.applyCurrentSource(1.);
      f.clearCurrent();
      restore();
      coupling.beginTrial();
      coupling.exchangeCandidate();
      pushDepositAll();
      coupling.completeCurrent();
      coupling.applyCurrentSource(-1.);
      coupling.exchangeCandidate();
      const auto r = coupling.localResidual();
      frc = r[0];
      const double total[3] = {r[1], r[2], r[3]};
      wrk = total[2] == 0. ? 0. : (total[1] > 0. ? total[0] / total[1] : total[0]);
      its = n + 1;
      if (its >= (total[2] > 0. ? 2u : 1u) && frc < 1e-13 && wrk < 1e-8) { converged = true; break; }
    }
    if (!converged) throw std::runtime_error("Conservative source coupling failed to converge");
    coupling.applyCurrentSource(1.);
    f.clearCurrent();
    restore();
    coupling.swapElectric();
    coupling.setReplay(true);
    pushDepositAll();
    coupling.completeCurrent();
    coupling.setReplay(false);
    coupling.finish(rf.step + 1, its, frc, wrk);
    coupling.swapElectric();
    rf.window->undoPreviewElectricBoundary(f);
    coupling.applyCurrentSource(-1.);
    rf.window->commitPreparedElectricBoundary(f);
    *iterations = its; *force = frc; *work = wrk;
  });
}

// Simulation::completeMovingInletLayer (simulation_moving_window.cpp:312-324)
// as the window's new-layer completion, over the coupling's particles.
namespace {
struct InletContext {
  RefCoupling* coupling = nullptr;
  FieldVector<Double> referenceOrigin{0.0};
  std::unique_ptr<ChargeMatchedInlet> inlet;
};
std::map<void*, std::unique_ptr<InletContext>> inletContexts;
void completeInlet(void* context, EBFieldGrid& f, std::size_t step, std::ptrdiff_t offset) {
  InletContext& c = *static_cast<InletContext*>(context);
  FieldVector<Double> origin = c.referenceOrigin;
  origin[2] = std::fma(static_cast<double>(offset), f.geometry().dz, c.referenceOrigin[2]);
  c.inlet->begin(origin);
  for (const auto& p : c.coupling->particles) c.inlet->particle(p);
  // Simulation::completeMovingInletLayer's pending visit (simulation_moving_window.cpp:317)
  if (c.coupling->pending) {
    const auto visit = [&](const RelativisticParticleSI& p) { c.inlet->particle(p); };
    if (c.coupling->quintic || c.coupling->binomial3) c.coupling->pending->visitNear(step, 8, visit);
    else c.coupling->pending->visitDue(step, visit);
  }
  c.inlet->complete(f, false);
}
}  // namespace

// The energy ledger's particle loop (simulation.cpp:3591-3625) over the
// particle list with the reference's BoostFrameTransform; the audit moment
// of ParticleEnergyMoment::sample (particle_energy_audit.h:20-40, a header
// that pulls in the HDF5 detectors) is written out here.
int ref_coupling_particle_energy(void* h, double timeBox, double* out) {
  return guarded([&] {
    RefCoupling& H = *RC(h);
    long double local[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (const RelativisticParticleSI& particle : H.particles) {
      const Double gammaBox = BoostFrameTransform::gammaFromProperVelocity(particle.properVelocity);
      const long double mass = static_cast<long double>(particle.mass);
      const long double c = static_cast<long double>(SI::c);
      const long double rest = mass * c * c;
      const long double u2 = static_cast<long double>(particle.properVelocity.norm2());
      local[0] += rest * u2 / static_cast<long double>(gammaBox + 1.0);
      local[1] += rest * static_cast<long double>(gammaBox);
      const long double represented = mass / static_cast<long double>(SI::electronMass);
      local[2] += represented;
      FieldVector<Double> properVelocityLab(0.0);
      H.frame.properVelocityBoxToLab(particle.properVelocity, properVelocityLab);
      const Double gammaLab = BoostFrameTransform::gammaFromProperVelocity(properVelocityLab);
      local[3] += represented * static_cast<long double>(gammaLab);
      local[4] += represented * static_cast<long double>(gammaBox);
      local[5] += represented * static_cast<long double>(H.frame.labZFromBoxZT(particle.position[2], timeBox));
      const long double restM = static_cast<long double>(particle.mass) * SI::c * SI::c;
      long double ul2 = 0;
      for (unsigned a = 0; a < 3; ++a) ul2 += static_cast<long double>(properVelocityLab[a]) * properVelocityLab[a];
      const long double gammaL = std::sqrt(1 + ul2);
      local[6] += restM * ul2 / (gammaL + 1);
      local[7] += restM * particle.properVelocity[2];
      local[8] += restM * properVelocityLab[2];
    }
    for (int k = 0; k < 9; ++k) out[k] = static_cast<double>(local[k]);
  });
}

int ref_window_set_charge_match(void* fields, void* coupling, const double* referenceOrigin) {
  return guarded([&] {
    RefFields& r = *R(fields);
    RefCoupling& H = *RC(coupling);
    std::unique_ptr<InletContext> c(new InletContext);
    c->coupling = &H;
    c->referenceOrigin = fv(referenceOrigin);
    c->inlet.reset(new ChargeMatchedInlet(r.f->geometry(), r.installed.cells[0], r.installed.cells[1], H.quintic,
                                          false, H.binomial3, H.smoothX));
    r.window->setNewLayerCompletion(&completeInlet, c.get(), true);
    inletContexts[fields] = std::move(c);
  });
}
int ref_window_begin_centered(void* h, std::size_t step, std::ptrdiff_t offset, const double* const* lower3,
                              const double* ezAtOffset) {
  return guarded([&] {
    RefFields& r = *R(h);
    r.staged.clear();
    for (unsigned a = 0; a < 3; ++a) r.stage(a, offset - 1, step, lower3[a]);
    r.stage(2, offset, step, ezAtOffset);
    r.step = step;
    r.window->beginStep(*r.f, step, offset);
  });
}

int ref_coupling_set_held(void* h, const int* frozen, const double* beta) {
  return guarded([&] {
    auto& P = RC(h)->particles;
    for (std::size_t i = 0; i < P.size(); ++i) {
      P[i].frozen = frozen[i] != 0;
      P[i].frozenBetaZ = beta[i];
    }
  });
}
int ref_coupling_get_frozen(void* h, int* frozen) {
  return guarded([&] {
    const auto& P = RC(h)->particles;
    for (std::size_t i = 0; i < P.size(); ++i) frozen[i] = P[i].frozen ? 1 : 0;
  });
}
int ref_coupling_set_activation(void* h, double labZ) {
  return guarded([&] { RC(h)->activationLabZ = labZ; });
}
int ref_coupling_set_hold_transverse(void* h, int hold) {
  return guarded([&] { RC(h)->holdTransverse = hold != 0; });
}

// The reference's PendingParticleQueue over the coupling's particles, as
// Simulation sets it up (simulation_moving_window.cpp:231-238).
int ref_coupling_set_pending(void* h, double referenceOriginZ, std::int64_t initialOffset, double dz,
                             double heldDistancePerBeta, std::uint64_t maximumStep, std::uint64_t initialStep) {
  return guarded([&] {
    RefCoupling& H = *RC(h);
    PendingWindowGeometry g;
    g.referenceOriginZ = referenceOriginZ;
    g.initialOffset = static_cast<std::ptrdiff_t>(initialOffset);
    g.dz = dz;
    g.heldDistancePerBeta = heldDistancePerBeta;
    g.maximumStep = static_cast<std::size_t>(maximumStep);
    H.pending.reset(new PendingParticleQueue(g, static_cast<std::size_t>(initialStep)));
    H.pending->extract(H.particles);
    H.pendingStep = static_cast<std::size_t>(initialStep);
  });
}
// Simulation::admitMovingWindowParticles (simulation_moving_window.cpp:275-278)
int ref_coupling_admit(void* h, std::uint64_t step, std::size_t* admitted) {
  return guarded([&] {
    RefCoupling& H = *RC(h);
    if (!H.pending) throw std::logic_error("No pending queue");
    *admitted = H.pending->admit(static_cast<std::size_t>(step), H.particles);
    H.pendingStep = static_cast<std::size_t>(step);
  });
}
int ref_coupling_pending_state(void* h, std::uint64_t* out3) {
  return guarded([&] {
    RefCoupling& H = *RC(h);
    if (!H.pending) throw std::logic_error("No pending queue");
    out3[0] = H.pending->size();
    out3[1] = H.pending->nextDueStep();
    out3[2] = H.pending->admittedCount();
  });
}
int ref_coupling_pending_near(void* h, std::uint64_t step, std::uint64_t lookAhead, std::size_t capacity, double* z,
                              std::uint64_t* id, std::size_t* n) {
  return guarded([&] {
    RefCoupling& H = *RC(h);
    if (!H.pending) throw std::logic_error("No pending queue");
    std::size_t count = 0;
    H.pending->visitNear(static_cast<std::size_t>(step), static_cast<std::size_t>(lookAhead),
                         [&](const RelativisticParticleSI& p) {
                           if (count < capacity) {
                             z[count] = p.position[2];
                             id[count] = p.id;
                           }
                           ++count;
                         });
    *n = count;
  });
}
int ref_repeat_held_addition(double initial, double increment, std::uint64_t count, double* out) {
  return guarded([&] { *out = repeatHeldAddition(initial, increment, static_cast<std::size_t>(count)); });
}

int ref_coupling_get_ids(void* h, std::uint64_t* id) {
  return guarded([&] {
    const auto& P = RC(h)->particles;
    for (std::size_t i = 0; i < P.size(); ++i) id[i] = P[i].id;
  });
}
std::size_t ref_coupling_carrier_count(void* h) { return RC(h)->carriers.size(); }
int ref_coupling_get_carriers(void* h, double* x3, double* v3, double* q, double* w, std::uint64_t* id) {
  return guarded([&] {
    const auto& C = RC(h)->carriers;
    for (std::size_t i = 0; i < C.size(); ++i) {
      for (int a = 0; a < 3; ++a) { x3[3 * i + a] = C[i].position[a]; v3[3 * i + a] = C[i].velocity[a]; }
      q[i] = C[i].charge;
      w[i] = C[i].currentWeight;
      id[i] = RC(h)->carrierIds[i];
    }
  });
}
int ref_coupling_counters(void* h, unsigned long long* counts, double* charges) {
  return guarded([&] {
    std::copy(RC(h)->counts, RC(h)->counts + 20, counts);
    std::copy(RC(h)->charges, RC(h)->charges + 20, charges);
  });
}
// The particle part of finishMovingWindowStep (simulation_moving_window.cpp:353-367).
int ref_coupling_finish_window(void* h, const double* newOrigin) {
  return guarded([&] {
    RefCoupling& H = *RC(h);
    const EBGridGeometry& g = H.fields->f->geometry();
    for (const auto& p : H.particles) {
      if (p.position[2] < newOrigin[2] && !p.frozen) throw std::runtime_error("Active particle fell behind moving inlet");
      // ownerRank (simulation.cpp:4068-4083) with one rank
      const Double coordinate = (p.position[2] - newOrigin[2]) / g.dz;
      const Double tolerance = 64.0 * std::numeric_limits<Double>::epsilon() * std::max(1.0, std::abs(coordinate));
      if (coordinate > static_cast<Double>(g.nz) + tolerance)
        throw std::runtime_error("Moving-window remap particle crossed a non-adjacent slab or outer boundary");
    }
    const double upper = std::fma(static_cast<double>(g.nz), g.dz, newOrigin[2]);
    std::size_t retained = 0;
    for (std::size_t n = 0; n < H.carriers.size(); ++n) {
      const auto& p = H.carriers[n];
      if (p.position[2] >= upper) {
        ++H.counts[19];
        H.charges[19] += p.charge * p.currentWeight;
      } else if (p.position[2] < newOrigin[2]) {
        throw std::runtime_error("Current carrier fell behind moving inlet");
      } else {
        H.carrierIds[retained] = H.carrierIds[n];
        H.carriers[retained++] = p;
      }
    }
    H.carriers.resize(retained);
    H.carrierIds.resize(retained);
  });
}
// ParticleCPMLRegion::firstEntry (kind 0), ParticleOpenBoundary::firstExit
// (kind 1) and currentDampingFactor (kind 2) at one grid origin.
int ref_boundary_query(void* h, const double* origin, int kind, const double* a, const double* b, double duration,
                       double* out6) {
  return guarded([&] {
    RefCoupling& H = *RC(h);
    const EBGridGeometry& g = H.fields->f->geometry();
    const FieldVector<Double> o = fv(origin);
    EBCPMLParameters p = H.fields->installed;
    for (int k = 0; k < 3; ++k) p.cells[k] = H.cpml.cells[k];
    ParticleCPMLRegion region(g, o, p);
    ParticleOpenBoundary outer(g, o, g, o, 0);
    if (kind == 2) {
      out6[0] = region.currentDampingFactor(fv(a), fv(b), duration);
      return;
    }
    ParticleBoundaryHit hit;
    const bool present = kind == 0 ? region.firstEntry(fv(a), fv(b), hit) : outer.firstExit(fv(a), fv(b), hit);
    out6[0] = present ? 1.0 : 0.0;
    out6[1] = hit.fraction;
    out6[2] = static_cast<double>(static_cast<int>(hit.face));
    for (int k = 0; k < 3; ++k) out6[3 + k] = hit.position[k];
  });
}

int ref_shape_weights(int shape, int smoothX, const double* geometry7, const double* origin, const double* begin,
                      const double* end, int* axis, long long* ijk, double* w, std::size_t capacity,
                      std::size_t* count, unsigned* segments) {
  return guarded([&] {
    EBGridGeometry g((std::size_t)geometry7[0], (std::size_t)geometry7[1], (std::size_t)geometry7[2], geometry7[3],
                     geometry7[4], geometry7[5], geometry7[6], EBMaxwellSolver::CowanZ);
    std::size_t n = 0;
    const auto sh = shape == 0 ? longitudinal_shape::Shape::QuinticCompensated
                               : (shape == 1 ? longitudinal_shape::Shape::LinearNotch : longitudinal_shape::Shape::LinearBinomial3);
    *segments = longitudinal_shape::weights(g, fv(origin), fv(begin), fv(end), [&](unsigned a, long long i, long long j, long long k, double v) {
      if (n >= capacity) throw std::length_error("weights capacity");
      axis[n] = (int)a; ijk[3 * n] = i; ijk[3 * n + 1] = j; ijk[3 * n + 2] = k; w[n] = v; ++n;
    }, sh, smoothX != 0);
    *count = n;
  });
}

int ref_momentum(double* u, const double* e, const double* b, double q, double m, double h, double g0,
                 unsigned* iterations) {
  return guarded([&] {
    FieldVector<Double> uu = fv(u);
    *iterations = source_coupling_math::momentum(uu, fv(e), fv(b), q, m, h, g0);
    for (int a = 0; a < 3; ++a) u[a] = uu[a];
  });
}

int ref_coupling_begin_fields(void* h, const double* origin) {
  return guarded([&] {
    RefCoupling& H = *RC(h);
    H.c->begin(*H.fields->f, fv(origin), H.particles, H.cpml.cells[0], H.cpml.cells[1]);
    H.c->beginTrial();
    H.c->exchangeCandidate();
  });
}

int ref_prescribed_box(void* h, const double* pos, double t, double* e, double* b) {
  return guarded([&] {
    FieldVector<Double> ee(0.0), bb(0.0);
    RC(h)->sources.addPrescribedBox(fv(pos), t, RC(h)->frame, ee, bb);
    for (int k = 0; k < 3; ++k) { e[k] = ee[k]; b[k] = bb[k]; }
  });
}

// Direct shape-level deposit with longitudinal_shape::Current (one rank).
int ref_current_deposit(void* fields, int shape, int smoothX, const double* origin, const double* a, const double* b,
                        double charge) {
  return guarded([&] {
    EBFieldGrid& f = *R(fields)->f;
    const auto sh = shape == 0 ? longitudinal_shape::Shape::QuinticCompensated
                               : (shape == 1 ? longitudinal_shape::Shape::LinearNotch : longitudinal_shape::Shape::LinearBinomial3);
    longitudinal_shape::Current cur(f, MPI_COMM_SELF, sh, 8, smoothX != 0);
    cur.deposit(fv(origin), fv(a), fv(b), charge);
  });
}

// Orbit gathers of source_coupling_math with an explicit old E (old = the
// fields' E at the call) for the binomial3-x and quintic paths.
int ref_binomial_xz_magnetic(void* fields, const double* origin, const double* point, double* out) {
  return guarded([&] {
    source_coupling_math::NativeHalo halo;
    const auto v = source_coupling_math::binomialXZMagneticSample(*R(fields)->f, fv(origin), fv(point), halo);
    for (int k = 0; k < 3; ++k) out[k] = v[k];
  });
}
int ref_magnetic_sample_open(void* fields, const double* origin, const double* point, double* out) {
  return guarded([&] {
    source_coupling_math::NativeHalo halo;
    const auto v = source_coupling_math::magneticSample(*R(fields)->f, fv(origin), fv(point), false, &halo, true).value;
    for (int k = 0; k < 3; ++k) out[k] = v[k];
  });
}

}  // extern "C"
