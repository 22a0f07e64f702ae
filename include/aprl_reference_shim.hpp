// aprl_reference_shim.hpp — the reference-side C++ binding of the MI355X
// library: what APRL would compile to call libaprl_gpu.so through the C-ABI
// of aprl_gpu.h with its own types. It includes the reference's headers
// (eb_field.h, eb_cpml.h, eb_particles.h, boostframe.h,
// moving_window_fields.h, eb_bunch.h), so it builds only inside the
// reference's source tree (or with it on the include path); it is not part
// of the library. Every call's status is rethrown as the reference's
// exception type with the library's message.
//
// What maps to what:
//   EBGridGeometry / EBFieldGrid / YeeComponent   -> FieldGrid (device arrays, host mirror on demand)
//   MovingWindowFields + MovingWindowReference    -> MovingWindow (the callback evaluated into planes)
//   std::vector<RelativisticParticleSI>           -> Particles (one device store, particles and carriers)
//   ConservativeSourceCoupling + solveConservativeSourceStep + commitConservativeElectric
//                                                 -> Coupling::step (one call per time step)
//   SIBunchPreprocessor (eb_bunch.h:62-109)       -> projectLabPlaneEventsToSnapshot,
//                                                    placeRelativeLabSnapshot, boostLabSnapshotToBoxTimeZero
//   verification::BlockedCowanReference           -> InletReference (prepared and kept on the device; a
//                                                    MovingWindow built on it fills its planes by a kernel)
#ifndef APRL_REFERENCE_SHIM_HPP
#define APRL_REFERENCE_SHIM_HPP

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "aprl_gpu.h"
#include "blocked_cowan_reference.h"
#include "boostframe.h"
#include "eb_bunch.h"
#include "eb_cpml.h"
#include "eb_field.h"
#include "eb_particles.h"
#include "moving_window_fields.h"

namespace aprl {
namespace gpu {

// aprl_status -> the reference's exception types (eb_field.cpp:880-941,
// simulation_source_coupling.cpp:54, moving_window_fields.cpp:208-209).
inline void check(aprl_status s) {
  if (s == APRL_OK) return;
  const std::string m = aprl_last_error();
  switch (s) {
    case APRL_INVALID_ARGUMENT: throw std::invalid_argument(m);
    case APRL_OUT_OF_RANGE: throw std::out_of_range(m);
    case APRL_LOGIC_ERROR: throw std::logic_error(m);
    default: throw std::runtime_error(m);
  }
}

inline aprl_grid_geometry geometryOf(const EBGridGeometry& g) {
  aprl_grid_geometry c;
  c.nx = g.nx;
  c.ny = g.ny;
  c.nz = g.nz;
  c.dx = g.dx;
  c.dy = g.dy;
  c.dz = g.dz;
  c.dt = g.dt;
  c.solver = g.solver == EBMaxwellSolver::CowanZ ? 1 : 0;
  return c;
}

inline aprl_cpml_parameters cpmlOf(const EBCPMLParameters& p) {
  aprl_cpml_parameters c;
  for (int a = 0; a < 3; ++a) c.cells[a] = p.cells[a];
  c.polynomial_order = p.polynomialOrder;
  c.target_reflection = p.targetReflection;
  c.kappa_max = p.kappaMax;
  c.alpha_fraction = p.alphaFraction;
  return c;
}

// BoostFrameTransform::setOriginsFromGamma's arguments (c0 = c).
inline aprl_boost boostOf(const BoostFrameTransform& f) {
  aprl_boost b;
  b.gamma = f.gamma();
  b.lab_time_origin = f.labTimeOrigin();
  b.lab_z_origin = f.labZOrigin();
  b.box_z_origin = f.boxZOrigin();
  return b;
}

// Component a (0..8 = Ex Ey Ez Bx By Bz Jx Jy Jz): nodes along x, y, z
// (eb_field.cpp:217-225; J is collocated with E).
inline void componentShape(const EBGridGeometry& g, int a, std::size_t n[3]) {
  const std::size_t cells[3] = {g.nx, g.ny, g.nz};
  const bool electric = a < 3 || a >= 6;
  const int axis = a % 3;
  for (int d = 0; d < 3; ++d) n[d] = cells[d] + ((d != axis) == electric ? 1 : 0);
}

// EBFieldGrid on the device. The reference's EBFieldGrid stays the host
// mirror that diagnostics, detectors and checkpoints read: download() fills
// it, upload() sends it (initial fields, restarts).
class FieldGrid {
 public:
  explicit FieldGrid(const EBGridGeometry& g, aprl_precision precision = APRL_FP64) : g_(g) {
    const aprl_grid_geometry c = geometryOf(g);
    check(aprl_fields_create(&c, precision, &f_));
  }
  ~FieldGrid() { aprl_fields_destroy(f_); }
  FieldGrid(const FieldGrid&) = delete;
  FieldGrid& operator=(const FieldGrid&) = delete;

  void upload(int a, const YeeComponent& c) { check(aprl_fields_upload(f_, a, c.data(), c.size())); }
  void download(int a, YeeComponent& c) const { check(aprl_fields_download(f_, a, c.data(), c.size())); }
  void upload(const EBFieldGrid& host) {
    const YeeComponent* c[9] = {&host.ex(), &host.ey(), &host.ez(), &host.bx(), &host.by(),
                                &host.bz(), &host.jx(), &host.jy(), &host.jz()};
    for (int a = 0; a < 9; ++a) upload(a, *c[a]);
  }
  void download(EBFieldGrid& host) const {
    YeeComponent* c[9] = {&host.ex(), &host.ey(), &host.ez(), &host.bx(), &host.by(),
                          &host.bz(), &host.jx(), &host.jy(), &host.jz()};
    for (int a = 0; a < 9; ++a) download(a, *c[a]);
  }
  // EBFieldGrid::setBoundary(EBConvolutionalPML) for a fixed grid
  void setCpml(const EBCPMLParameters& p, std::size_t globalNz, std::size_t localZOffset, bool lowerZPhysical,
               bool upperZPhysical) {
    const aprl_cpml_parameters c = cpmlOf(p);
    check(aprl_fields_set_cpml(f_, &c, globalNz, localZOffset, lowerZPhysical ? 1 : 0, upperZPhysical ? 1 : 0));
  }
  void clearCurrent() { check(aprl_fields_clear_current(f_)); }
  void advanceMagnetic() { check(aprl_fields_advance_magnetic(f_, nullptr)); }
  void advanceElectric() { check(aprl_fields_advance_electric(f_)); }
  void synchronize() { check(aprl_fields_synchronize(f_)); }

  aprl_fields* handle() const { return f_; }
  const EBGridGeometry& geometry() const { return g_; }

 private:
  EBGridGeometry g_;
  aprl_fields* f_ = nullptr;
};

// verification::BlockedCowanReference (blocked_cowan_reference.h:68-129)
// for one rank, prepared on the device at construction and kept in device
// memory: no work directory, no block files, no communicator. field() reads
// a host copy of the component made on first use; fill() writes E/B of a
// grid of the same x-y geometry from it; a MovingWindow built on it takes
// movingWindowFieldValue's planes from it without a host callback. The
// open-wall transverse reference is not built (openTransverse rejected).
class InletReference {
 public:
  InletReference(const EBGridGeometry& globalGeometry, const FieldVector<Double>& globalOrigin,
                 const std::vector<RelativisticParticleSI>& particles, Double beta, const std::size_t guards[3],
                 bool longitudinalQuintic = false, bool linearNotch = false, bool linearBinomial3 = false,
                 bool transverseXBinomial3 = false, bool openTransverse = false)
      : g_(globalGeometry) {
    if (static_cast<unsigned>(longitudinalQuintic) + static_cast<unsigned>(linearNotch) +
            static_cast<unsigned>(linearBinomial3) > 1)
      throw std::invalid_argument("Ambiguous reference source shape");
    if (linearNotch) throw std::invalid_argument("The retired linear-notch reference source is not built");
    if (openTransverse) throw std::invalid_argument("The open-wall transverse reference is not built");
    const std::size_t n = particles.size();
    std::vector<double> x(3 * n), q(n);
    for (std::size_t p = 0; p < n; ++p) {
      for (int a = 0; a < 3; ++a) x[3 * p + a] = particles[p].position[a];
      q[p] = particles[p].charge;
    }
    const aprl_grid_geometry c = geometryOf(globalGeometry);
    const double o[3] = {globalOrigin[0], globalOrigin[1], globalOrigin[2]};
    const int shape = longitudinalQuintic ? APRL_SHAPE_QUINTIC_COMPENSATED
                                          : (linearBinomial3 ? APRL_SHAPE_LINEAR_BINOMIAL3 : -1);
    aprl_inlet_reference_report r;
    check(aprl_inlet_reference_create(&c, o, n, x.data(), q.data(), beta, guards, shape,
                                      transverseXBinomial3 ? 1 : 0, &r_, &r));
    report_.complete = true;
    report_.particles = n;
    for (int a = 0; a < 3; ++a) report_.interiorNodes[a] = r.interior_nodes[a];
    report_.beta = beta;
    report_.sourceCharge = r.source_charge;
    report_.depositedCharge = r.deposited_charge;
    report_.minimumPivot = r.minimum_pivot;
    report_.minimumModeEigenvalue = r.minimum_mode_eigenvalue;
  }
  ~InletReference() { aprl_inlet_reference_destroy(r_); }
  InletReference(const InletReference&) = delete;
  InletReference& operator=(const InletReference&) = delete;

  // native E(0), B(-dt/2) at global native indices (component order Ex..Bz)
  Double field(unsigned component, std::size_t i, std::size_t j, std::size_t globalK) {
    std::size_t n[3] = {0, 0, 0};
    if (component < 6) componentShape(g_, static_cast<int>(component), n);
    if (component >= 6 || i >= n[0] || j >= n[1] || globalK >= n[2])
      throw std::out_of_range("Blocked field query outside native global Yee grid");
    std::vector<double>& h = host_[component];
    if (h.empty()) {
      h.resize(n[0] * n[1] * n[2]);
      check(aprl_inlet_reference_download(r_, static_cast<int>(component), h.data(), h.size()));
    }
    return h[(globalK * n[1] + j) * n[0] + i];
  }
  // BlockedCowanReference::fill: E/B of a grid with the same x-y geometry
  // from global plane globalZOffset on; J is left alone.
  void fill(FieldGrid& fields, std::size_t globalZOffset = 0) {
    // blocked_cowan_reference.cpp:958-963
    const EBGridGeometry& w = fields.geometry();
    if (w.nx != g_.nx || w.ny != g_.ny || w.dx != g_.dx || w.dy != g_.dy || w.dz != g_.dz || w.dt != g_.dt ||
        w.solver != g_.solver || globalZOffset > g_.nz || w.nz > g_.nz - globalZOffset)
      throw std::invalid_argument("Blocked reference fill geometry/offset mismatch");
    for (unsigned a = 0; a < 6; ++a) {
      std::size_t n[3], m[3];
      componentShape(g_, static_cast<int>(a), n);
      componentShape(w, static_cast<int>(a), m);
      field(a, 0, 0, 0);
      const std::vector<double>& h = host_[a];
      check(aprl_fields_upload(fields.handle(), static_cast<int>(a), h.data() + globalZOffset * n[0] * n[1],
                               m[0] * m[1] * m[2]));
    }
  }
  void clearCache() {
    for (auto& h : host_) std::vector<double>().swap(h);
  }
  const verification::BlockedCowanReferenceReport& report() const { return report_; }
  aprl_inlet_reference* handle() const { return r_; }

 private:
  EBGridGeometry g_;
  aprl_inlet_reference* r_ = nullptr;
  verification::BlockedCowanReferenceReport report_;
  std::vector<double> host_[6];
};

// MovingWindowFields on the device. The reference callback stays on the host
// (moving_window_fields.h:15-20): every step it is evaluated into the
// planes the device needs, reference(a, i, j, k, step) over component a's
// x-y nodes (fixed-world k, not window-relative).
class MovingWindow {
 public:
  MovingWindow(FieldGrid& grid, const EBCPMLParameters& cpml, MovingWindowReference reference)
      : grid_(grid), ref_(reference) {
    const aprl_cpml_parameters c = cpmlOf(cpml);
    check(aprl_window_install(grid.handle(), &c));
  }
  // the production reference (movingWindowFieldValue over a
  // BlockedCowanReference translating at translationBeta,
  // simulation_moving_window.cpp:280-294): planes filled on the device
  MovingWindow(FieldGrid& grid, const EBCPMLParameters& cpml, InletReference& reference, Double translationBeta)
      : grid_(grid), ref_{} {
    const aprl_cpml_parameters c = cpmlOf(cpml);
    check(aprl_window_install(grid.handle(), &c));
    check(aprl_window_set_reference(grid.handle(), reference.handle(), translationBeta));
  }
  // MovingWindowFields::beginStep: lower E halo planes at offset - 1.
  void beginStep(std::size_t step, std::ptrdiff_t offset) {
    const double* lower[3] = {plane(0, offset - 1, step), plane(1, offset - 1, step), plane(2, offset - 1, step)};
    check(aprl_window_begin_step(grid_.handle(), step, offset, lower));
  }
  // the inlet overwrite values reference(a, i, j, offset, step + 1) of Ex, Ey
  void setInlet(std::size_t step, std::ptrdiff_t offset) {
    const double* ex = plane(0, offset, step + 1);
    const double* ey = plane(1, offset, step + 1);
    check(aprl_window_set_inlet(grid_.handle(), ex, ey));
  }
  // MovingWindowFields::shift to (newStep, newOffset): six inserted planes
  void shift(std::size_t newStep, std::ptrdiff_t newOffset) {
    const double* planes[6];
    for (unsigned a = 0; a < 6; ++a) planes[a] = plane(a, newOffset, newStep);
    check(aprl_window_shift(grid_.handle(), newStep, newOffset, planes));
  }
  void enableEnergyLedger(bool enabled) { check(aprl_window_enable_energy_ledger(grid_.handle(), enabled ? 1 : 0)); }
  void finishEnergyLedgerStep() { check(aprl_window_finish_energy_step(grid_.handle())); }

 private:
  // one plane per component, reused every call (a null callback is the zero
  // reference and moves no data)
  const double* plane(unsigned a, std::ptrdiff_t k, std::size_t step) {
    if (!ref_.value) return nullptr;
    std::size_t n[3];
    componentShape(grid_.geometry(), static_cast<int>(a), n);
    std::vector<double>& p = planes_[a];
    p.resize(n[0] * n[1]);
    for (std::size_t j = 0; j < n[1]; ++j)
      for (std::size_t i = 0; i < n[0]; ++i) p[j * n[0] + i] = ref_.value(ref_.context, a, i, j, k, step);
    return p.data();
  }

  FieldGrid& grid_;
  MovingWindowReference ref_;
  std::vector<double> planes_[6];
};

// std::vector<RelativisticParticleSI> on the device: one store for particles
// (held ones included) and CPML carriers, in tile-sorted order with the ids
// travelling with the entries.
class Particles {
 public:
  Particles(FieldGrid& grid, std::size_t capacity) { check(aprl_particles_create(grid.handle(), capacity, &p_)); }
  ~Particles() { aprl_particles_destroy(p_); }
  Particles(const Particles&) = delete;
  Particles& operator=(const Particles&) = delete;

  // particles_ with their frozen flags and held speeds (weight: the
  // diagnostic weight, eb_particles.h:18)
  void upload(const std::vector<RelativisticParticleSI>& particles) {
    const std::size_t n = particles.size();
    std::vector<double> x(3 * n), u(3 * n), q(n), m(n), w(n), beta(n);
    std::vector<uint64_t> id(n);
    std::vector<int> frozen(n);
    for (std::size_t p = 0; p < n; ++p) {
      const RelativisticParticleSI& s = particles[p];
      for (int a = 0; a < 3; ++a) {
        x[3 * p + a] = s.position[a];
        u[3 * p + a] = s.properVelocity[a];
      }
      q[p] = s.charge;
      m[p] = s.mass;
      w[p] = s.weight;
      id[p] = s.id;
      frozen[p] = s.frozen ? 1 : 0;
      beta[p] = s.frozen ? s.frozenBetaZ : 0.0;
    }
    check(aprl_particles_upload(p_, n, x.data(), u.data(), q.data(), m.data(), w.data(), id.data()));
    check(aprl_particles_set_held(p_, frozen.data(), beta.data()));
  }
  // the store split into particles_ (active and held) and carriers (state
  // 1: position, u and the current weight in weight)
  void download(std::vector<RelativisticParticleSI>& particles, std::vector<RelativisticParticleSI>* carriers) {
    const std::size_t n = aprl_particles_count(p_);
    std::vector<double> x(3 * n), u(3 * n), q(n), m(n), w(n), beta(n);
    std::vector<uint64_t> id(n);
    std::vector<int> state(n);
    check(aprl_particles_download(p_, x.data(), u.data(), q.data(), m.data(), w.data(), id.data()));
    check(aprl_particles_download_states(p_, state.data()));
    check(aprl_particles_download_held(p_, beta.data()));
    particles.clear();
    if (carriers) carriers->clear();
    for (std::size_t p = 0; p < n; ++p) {
      RelativisticParticleSI s;
      for (int a = 0; a < 3; ++a) {
        s.position[a] = x[3 * p + a];
        s.properVelocity[a] = u[3 * p + a];
      }
      s.charge = q[p];
      s.mass = m[p];
      s.weight = w[p];
      s.id = id[p];
      s.frozen = state[p] == 2;
      s.frozenBetaZ = beta[p];
      if (state[p] == 1) {
        if (carriers) carriers->push_back(s);
      } else {
        particles.push_back(s);
      }
    }
  }
  std::size_t size() const { return aprl_particles_count(p_); }
  aprl_particles* handle() const { return p_; }

 private:
  aprl_particles* p_ = nullptr;
};

// ConservativeSourceCoupling and the Picard driver around it: one step()
// replaces solveConservativeSourceStep + commitConservativeElectric
// (simulation_source_coupling.cpp:16-65) for every particle at once.
class Coupling {
 public:
  Coupling(FieldGrid& grid, bool quintic, bool transverseXBinomial3) {
    check(aprl_coupling_create(grid.handle(), quintic ? APRL_SHAPE_QUINTIC_COMPENSATED : APRL_SHAPE_LINEAR_BINOMIAL3,
                               transverseXBinomial3 ? 1 : 0, &c_));
  }
  ~Coupling() { aprl_coupling_destroy(c_); }
  Coupling(const Coupling&) = delete;
  Coupling& operator=(const Coupling&) = delete;

  void setFrame(const BoostFrameTransform& frame) {
    const aprl_boost b = boostOf(frame);
    check(aprl_coupling_set_frame(c_, &b));
  }
  void addPrescribedLabMagnet(const aprl_magnet& magnet) { check(aprl_coupling_add_magnet(c_, &magnet)); }
  void setParticleBoundary(const std::size_t cpmlCells[3], Double exitLabZ) {
    check(aprl_coupling_set_particle_boundary(c_, cpmlCells, exitLabZ));
  }
  void setActivationPlane(Double labZ) { check(aprl_coupling_set_activation(c_, labZ)); }
  void setHoldTransverse(bool hold) { check(aprl_coupling_set_hold_transverse(c_, hold ? 1 : 0)); }
  aprl_step_report step(Particles& particles, const FieldVector<Double>& origin, Double time, unsigned substeps) {
    const double o[3] = {origin[0], origin[1], origin[2]};
    aprl_step_report r;
    check(aprl_coupling_step(c_, particles.handle(), o, time, substeps, &r));
    return r;
  }
  aprl_coupling* handle() const { return c_; }

 private:
  aprl_coupling* c_ = nullptr;
};

// SIBunchPreprocessor's functions (eb_bunch.h:62-109) for one rank, on the
// device: the vector goes to a device bunch and comes back transformed.
namespace detail {
class DeviceBunch {
 public:
  explicit DeviceBunch(const std::vector<RelativisticParticleSI>& particles) {
    const std::size_t n = particles.size();
    check(aprl_bunch_create(n, &b_));
    std::vector<double> x(3 * n), u(3 * n), q(n), m(n), w(n);
    std::vector<uint64_t> id(n);
    for (std::size_t p = 0; p < n; ++p) {
      for (int a = 0; a < 3; ++a) {
        x[3 * p + a] = particles[p].position[a];
        u[3 * p + a] = particles[p].properVelocity[a];
      }
      q[p] = particles[p].charge;
      m[p] = particles[p].mass;
      w[p] = particles[p].weight;
      id[p] = particles[p].id;
    }
    check(aprl_bunch_upload(b_, n, x.data(), u.data(), q.data(), m.data(), w.data(), id.data()));
  }
  ~DeviceBunch() { aprl_bunch_destroy(b_); }
  DeviceBunch(const DeviceBunch&) = delete;
  DeviceBunch& operator=(const DeviceBunch&) = delete;
  // positions, proper velocities and held speeds back into the vector
  void storeInto(std::vector<RelativisticParticleSI>& particles) const {
    const std::size_t n = particles.size();
    std::vector<double> x(3 * n), u(3 * n), beta(n);
    check(aprl_bunch_download(b_, x.data(), u.data(), nullptr, nullptr, nullptr, nullptr, beta.data()));
    for (std::size_t p = 0; p < n; ++p) {
      for (int a = 0; a < 3; ++a) {
        particles[p].position[a] = x[3 * p + a];
        particles[p].properVelocity[a] = u[3 * p + a];
      }
      if (particles[p].frozen) particles[p].frozenBetaZ = beta[p];
    }
  }
  aprl_bunch* handle() const { return b_; }

 private:
  aprl_bunch* b_ = nullptr;
};
}  // namespace detail

inline SILabPlaneProjectionReport projectLabPlaneEventsToSnapshot(std::vector<RelativisticParticleSI>& particles,
                                                                  Double inputPlaneLab, Double referencePositionLab,
                                                                  Double longitudinalOffsetLab,
                                                                  bool holdTransverseUntilActivation = false) {
  detail::DeviceBunch b(particles);
  aprl_lab_plane_projection r;
  check(aprl_bunch_project_lab_plane_events(b.handle(), inputPlaneLab, referencePositionLab, longitudinalOffsetLab,
                                            holdTransverseUntilActivation ? 1 : 0, &r));
  b.storeInto(particles);
  SILabPlaneProjectionReport out;
  out.inputPlaneLab = r.input_plane_lab;
  out.referencePositionLab = r.reference_position_lab;
  out.minimumForwardDistance = r.minimum_forward_distance;
  out.maximumForwardDistance = r.maximum_forward_distance;
  out.recommendedMinimumReferencePosition = r.recommended_minimum_reference_position;
  out.referenceTimeOffsetLab = r.reference_time_offset_lab;
  out.meanLongitudinalBeta = r.mean_longitudinal_beta;
  out.particles = r.particles;
  return out;
}

inline SIBunchPlacementReport placeRelativeLabSnapshot(std::vector<RelativisticParticleSI>& particles,
                                                       const SIBunchPlacement& placement) {
  detail::DeviceBunch b(particles);
  aprl_bunch_placement r;
  check(aprl_bunch_place_relative(b.handle(), placement.hasFirstInteractionEntrance ? 1 : 0,
                                  placement.firstInteractionEntranceLab, placement.referencePositionLab,
                                  placement.recommendationMarginLab, &r));
  b.storeInto(particles);
  SIBunchPlacementReport out;
  out.relativeTailLab = r.relative_tail_lab;
  out.relativeHeadLab = r.relative_head_lab;
  out.referencePositionLab = r.reference_position_lab;
  out.headAfterLab = r.head_after_lab;
  out.actualHeadDistance = r.actual_head_distance;
  out.recommendedMaximumReferencePosition = r.recommended_maximum_reference_position;
  out.particles = r.particles;
  return out;
}

inline SIBunchBoostReport boostLabSnapshotToBoxTimeZero(std::vector<RelativisticParticleSI>& particles,
                                                        const BoostFrameTransform& frame, Double snapshotTimeLab = 0.0,
                                                        bool holdTransverseUntilActivation = false,
                                                        Double inputPlaneLab = 0.0) {
  detail::DeviceBunch b(particles);
  const aprl_boost f = boostOf(frame);
  aprl_bunch_boost_report r;
  check(aprl_bunch_boost_to_box_time_zero(b.handle(), &f, snapshotTimeLab, holdTransverseUntilActivation ? 1 : 0,
                                          inputPlaneLab, &r));
  b.storeInto(particles);
  SIBunchBoostReport out;
  out.earliestLabEventTime = r.earliest_lab_event_time;
  out.latestLabEventTime = r.latest_lab_event_time;
  out.maximumAbsoluteDriftTime = r.maximum_absolute_drift_time;
  out.minimumLabGamma = r.minimum_lab_gamma;
  out.maximumLabGamma = r.maximum_lab_gamma;
  out.maximumRelativeMomentumRoundTripError = r.maximum_relative_momentum_round_trip_error;
  out.maximumRelativeGammaRoundTripError = r.maximum_relative_gamma_round_trip_error;
  return out;
}

}  // namespace gpu
}  // namespace aprl

#endif  // APRL_REFERENCE_SHIM_HPP
