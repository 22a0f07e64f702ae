// C-ABI entry points for the particle store and the conservative-source
// coupling (declared in include/aprl_gpu.h).
#include <cmath>
#include <memory>

#include "diagnostics.hpp"
#include "handles.hpp"
#include "inlet.hpp"

namespace aprl_gpu {
void couplingStep(CouplingState& C, ParticleStore& P, const double origin[3], double time, unsigned substeps,
                  aprl_step_report* report);
}  // namespace aprl_gpu

using namespace aprl_gpu;

namespace {
// SIMagneticElement::prepare, eb_sources.cpp:320-344 (host, same libm).
MagnetDev prepareMagnet(const aprl_magnet& in) {
  MagnetDev m{};
  m.type = in.type;
  for (int a = 0; a < 3; ++a) m.center[a] = in.center[a];
  m.length = in.length;
  m.period = in.period;
  m.peak = in.peak_field;
  m.angle = in.polarization_angle;
  m.gaussianFringe = in.gaussian_fringe;
  m.cutoff = in.fringe_relative_cutoff;
  m.gradient = in.quadrupole_gradient;
  m.endRampPeriods = in.end_ramp_periods;
  m.phase = in.undulator_phase;
  if (m.type < 0 || m.type > 2) throw std::invalid_argument("Unknown magnetic element type");
  for (double c : m.center)
    if (!std::isfinite(c)) throw std::invalid_argument("Invalid SI magnetic element");
  if (!(m.length >= 0.0) || !std::isfinite(m.length) || !std::isfinite(m.peak) || !std::isfinite(m.angle))
    throw std::invalid_argument("Invalid SI magnetic element");
  if (m.type == 0 && (!(m.period > 0.0) || !std::isfinite(m.period)))
    throw std::invalid_argument("Undulator period must be positive");
  if (m.type == 0 && m.gaussianFringe && (!(m.cutoff > 0.0) || !(m.cutoff < 1.0) || !std::isfinite(m.cutoff)))
    throw std::invalid_argument("Undulator fringe_relative_cutoff must be in (0,1)");
  if (m.type == 1 && !(m.length > 0.0)) throw std::invalid_argument("Dipole length must be positive");
  if (m.type == 2 && (!(m.length > 0.0) || !std::isfinite(m.gradient)))
    throw std::invalid_argument("Quadrupole length/gradient is invalid");
  if (!std::isfinite(m.endRampPeriods) || m.endRampPeriods < 0.0 ||
      (m.endRampPeriods > 0.0 && (m.type != 0 || m.gaussianFringe || 2.0 * m.endRampPeriods * m.period > m.length)))
    throw std::invalid_argument("Internal undulator end ramp requires non-Gaussian ends and fits inside length");
  if (!std::isfinite(m.phase) || (m.phase != 0.0 && (m.type != 0 || m.endRampPeriods <= 0.0)))
    throw std::invalid_argument("Undulator phase requires internal end ramps");
  const double twoPi = 2.0 * 3.141592653589793238462643383279502884;
  m.waveNumber = m.type == 0 ? twoPi / m.period : 0.0;
  m.fringeExtent = m.type == 0 && m.gaussianFringe ? std::sqrt(-2.0 * std::log(m.cutoff)) / m.waveNumber : 0.0;
  m.cosA = std::cos(m.angle);
  m.sinA = std::sin(m.angle);
  return m;
}
}  // namespace

extern "C" {

aprl_status aprl_particles_create(aprl_fields* fields, size_t capacity, aprl_particles** out) {
  return guarded([&] {
    if (!fields || !out) throw std::invalid_argument("Null field grid or output handle");
    std::unique_ptr<aprl_particles> p(new aprl_particles);
    p->fields = fields;
    if (fields->fp32) reserveSplitParticles(p->sp, std::max<size_t>(capacity, 1));
    else reserveParticles(p->p, std::max<size_t>(capacity, 1));
    *out = p.release();
  });
}

void aprl_particles_destroy(aprl_particles* particles) {
  if (!particles) return;
  (void)hipStreamSynchronize(particles->stream());
  delete particles;
}

size_t aprl_particles_count(const aprl_particles* particles) {
  if (!particles) return 0;
  try {
    return particles->fields->fp32 ? liveEntries(particles->sp.count, particles->sp.removed, particles->stream())
                                   : liveEntries(particles->p.count, particles->p.removed, particles->stream());
  } catch (...) {
    return 0;
  }
}

aprl_status aprl_particles_upload(aprl_particles* particles, size_t n, const double* x3, const double* u3,
                                  const double* q, const double* m, const double* w, const uint64_t* id) {
  return guarded([&] {
    if (!particles) throw std::invalid_argument("Null particle store");
    if (n && (!x3 || !u3 || !q || !m || !w || !id)) throw std::invalid_argument("Null particle array");
    hipCheck(hipStreamSynchronize(particles->stream()), "hipStreamSynchronize");
    int* state = nullptr;
    unsigned* removed = nullptr;
    if (particles->fields->fp32) {
      uploadSplitParticles(particles->sp, particles->fields->s32, n, x3, u3, q, m, w, id);
      state = particles->sp.state.get();
      removed = particles->sp.removed.get();
    } else {
      ParticleStore& P = particles->p;
      reserveParticles(P, n);
      std::vector<double> col(n);
      for (int a = 0; a < 3; ++a) {
        for (size_t i = 0; i < n; ++i) col[i] = x3[3 * i + a];
        hipCheck(hipMemcpy(P.s[a].get(), col.data(), n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        for (size_t i = 0; i < n; ++i) col[i] = u3[3 * i + a];
        hipCheck(hipMemcpy(P.s[3 + a].get(), col.data(), n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
      }
      const double* scalars[3] = {q, m, w};
      for (int a = 0; a < 3; ++a)
        hipCheck(hipMemcpy(P.s[6 + a].get(), scalars[a], n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
      hipCheck(hipMemcpy(P.id.get(), id, n * sizeof(uint64_t), hipMemcpyHostToDevice), "hipMemcpy");
      P.count = n;
      ++P.generation;
      state = P.state.get();
      removed = P.removed.get();
    }
    // every uploaded entry is an active particle
    if (n) {
      hipCheck(hipMemset(state, 0, n * sizeof(int)), "hipMemset");
      if (particles->fields->fp32) hipCheck(hipMemset(particles->sp.beta.get(), 0, n * sizeof(float)), "hipMemset");
      else hipCheck(hipMemset(particles->p.beta.get(), 0, n * sizeof(double)), "hipMemset");
    }
    hipCheck(hipMemset(removed, 0, sizeof(unsigned)), "hipMemset");
  });
}

namespace {
// Entry states of the whole store (live and not yet compacted entries).
std::vector<int> storeStates(const aprl_particles* particles, std::size_t n) {
  std::vector<int> st(n);
  const int* src = particles->fields->fp32 ? particles->sp.state.get() : particles->p.state.get();
  if (n) hipCheck(hipMemcpy(st.data(), src, n * sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy");
  return st;
}

void downloadAll(aprl_particles* particles, double* x3, double* u3, double* q, double* m, double* w, uint64_t* id) {
  if (particles->fields->fp32) {
    downloadSplitParticles(particles->sp, particles->fields->s32, x3, u3, q, m, w, id);
    return;
  }
  ParticleStore& P = particles->p;
  const size_t n = P.count;
  std::vector<double> col(n);
  for (int a = 0; a < 3; ++a) {
    if (x3) {
      hipCheck(hipMemcpy(col.data(), P.s[a].get(), n * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
      for (size_t i = 0; i < n; ++i) x3[3 * i + a] = col[i];
    }
    if (u3) {
      hipCheck(hipMemcpy(col.data(), P.s[3 + a].get(), n * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
      for (size_t i = 0; i < n; ++i) u3[3 * i + a] = col[i];
    }
  }
  double* scalars[3] = {q, m, w};
  for (int a = 0; a < 3; ++a)
    if (scalars[a])
      hipCheck(hipMemcpy(scalars[a], P.s[6 + a].get(), n * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
  if (id) hipCheck(hipMemcpy(id, P.id.get(), n * sizeof(uint64_t), hipMemcpyDeviceToHost), "hipMemcpy");
}
}  // namespace

aprl_status aprl_particles_download(aprl_particles* particles, double* x3, double* u3, double* q, double* m,
                                    double* w, uint64_t* id) {
  return guarded([&] {
    if (!particles) throw std::invalid_argument("Null particle store");
    hipCheck(hipStreamSynchronize(particles->stream()), "hipStreamSynchronize");
    const size_t n = particles->fields->fp32 ? particles->sp.count : particles->p.count;
    const size_t live = aprl_particles_count(particles);
    if (live == n) {
      downloadAll(particles, x3, u3, q, m, w, id);
      return;
    }
    // entries removed by the last step are skipped (the next sort drops them)
    std::vector<double> X(x3 ? 3 * n : 0), U(u3 ? 3 * n : 0), Q(q ? n : 0), M(m ? n : 0), W(w ? n : 0);
    std::vector<uint64_t> I(id ? n : 0);
    downloadAll(particles, x3 ? X.data() : nullptr, u3 ? U.data() : nullptr, q ? Q.data() : nullptr,
                m ? M.data() : nullptr, w ? W.data() : nullptr, id ? I.data() : nullptr);
    const std::vector<int> st = storeStates(particles, n);
    size_t k = 0;
    for (size_t i = 0; i < n; ++i) {
      if (st[i] < 0) continue;
      for (int a = 0; a < 3; ++a) {
        if (x3) x3[3 * k + a] = X[3 * i + a];
        if (u3) u3[3 * k + a] = U[3 * i + a];
      }
      if (q) q[k] = Q[i];
      if (m) m[k] = M[i];
      if (w) w[k] = W[i];
      if (id) id[k] = I[i];
      ++k;
    }
  });
}

aprl_status aprl_particles_set_held(aprl_particles* particles, const int* frozen, const double* beta) {
  return guarded([&] {
    if (!particles || !frozen || !beta) throw std::invalid_argument("Null particle store or held arrays");
    hipCheck(hipStreamSynchronize(particles->stream()), "hipStreamSynchronize");
    const size_t n = particles->fields->fp32 ? particles->sp.count : particles->p.count;
    if (aprl_particles_count(particles) != n)
      throw std::logic_error("Held states are set on a store without removed entries (after an upload)");
    std::vector<int> st = storeStates(particles, n);
    for (size_t i = 0; i < n; ++i) {
      if (st[i] == kStateCarrier) {
        if (frozen[i]) throw std::invalid_argument("A CPML carrier cannot be held");
        continue;
      }
      if (frozen[i] && !(std::isfinite(beta[i]) && beta[i] >= 0.0 && beta[i] < 1.0))
        throw std::invalid_argument("Held beta must be finite in [0, 1)");
      st[i] = frozen[i] ? kStateHeld : kStateActive;
    }
    int* dst = particles->fields->fp32 ? particles->sp.state.get() : particles->p.state.get();
    if (n) hipCheck(hipMemcpy(dst, st.data(), n * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
    if (particles->fields->fp32) {
      std::vector<float> b(n);
      for (size_t i = 0; i < n; ++i) b[i] = static_cast<float>(beta[i]);
      if (n) hipCheck(hipMemcpy(particles->sp.beta.get(), b.data(), n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
    } else if (n) {
      hipCheck(hipMemcpy(particles->p.beta.get(), beta, n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    }
  });
}

aprl_status aprl_coupling_set_activation(aprl_coupling* coupling, double activation_lab_z) {
  return guarded([&] {
    if (!coupling) throw std::invalid_argument("Null coupling");
    coupling->c.hasActivation = !std::isnan(activation_lab_z);
    coupling->c.activationLabZ = activation_lab_z;
  });
}

aprl_status aprl_coupling_set_hold_transverse(aprl_coupling* coupling, int hold) {
  return guarded([&] {
    if (!coupling) throw std::invalid_argument("Null coupling");
    coupling->c.holdTransverse = hold != 0;
  });
}

aprl_status aprl_particles_download_states(aprl_particles* particles, int* state) {
  return guarded([&] {
    if (!particles || !state) throw std::invalid_argument("Null particle store or state array");
    hipCheck(hipStreamSynchronize(particles->stream()), "hipStreamSynchronize");
    const size_t n = particles->fields->fp32 ? particles->sp.count : particles->p.count;
    const std::vector<int> st = storeStates(particles, n);
    size_t k = 0;
    for (size_t i = 0; i < n; ++i)
      if (st[i] >= 0) state[k++] = st[i];
  });
}

aprl_status aprl_particles_download_held(aprl_particles* particles, double* beta) {
  return guarded([&] {
    if (!particles || !beta) throw std::invalid_argument("Null particle store or held-speed array");
    hipCheck(hipStreamSynchronize(particles->stream()), "hipStreamSynchronize");
    const size_t n = particles->fields->fp32 ? particles->sp.count : particles->p.count;
    const std::vector<int> st = storeStates(particles, n);
    std::vector<double> b(n);
    if (particles->fields->fp32) {
      std::vector<float> f(n);
      if (n) hipCheck(hipMemcpy(f.data(), particles->sp.beta.get(), n * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
      for (size_t i = 0; i < n; ++i) b[i] = static_cast<double>(f[i]);
    } else if (n) {
      hipCheck(hipMemcpy(b.data(), particles->p.beta.get(), n * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    }
    size_t k = 0;
    for (size_t i = 0; i < n; ++i)
      if (st[i] >= 0) beta[k++] = st[i] == kStateHeld ? b[i] : 0.0;
  });
}

aprl_status aprl_coupling_set_reference_current(aprl_coupling* coupling, const double* table, size_t rows,
                                               size_t count, const size_t* guards, double translation_beta) {
  return guarded([&] {
    if (!coupling) throw std::invalid_argument("Null coupling");
    CouplingState& C = coupling->c;
    if (!table) {
      C.refCurrent = false;
      return;
    }
    if (!guards) throw std::invalid_argument("Null reference-current guards");
    const aprl_grid_geometry& g = coupling->fields->fp32 ? coupling->fields->s32.g : coupling->fields->s.g;
    if (!guards[0] || !guards[1]) throw std::invalid_argument("Reference surface current requires transverse-only CPML");
    if (2 * guards[0] >= g.nx || 2 * guards[1] >= g.ny) throw std::invalid_argument("Reference-current guards leave no aperture");
    // TransverseReferenceCurrent points() (transverse_reference_current.cpp:18-26)
    std::vector<int> ij;
    const size_t x0 = guards[0], x1 = g.nx - guards[0], y0 = guards[1], y1 = g.ny - guards[1];
    for (size_t i = x0; i <= x1; ++i)
      for (size_t j : {y0, y1}) { ij.push_back(static_cast<int>(i)); ij.push_back(static_cast<int>(j)); }
    for (size_t j = y0 + 1; j < y1; ++j)
      for (size_t i : {x0, x1}) { ij.push_back(static_cast<int>(i)); ij.push_back(static_cast<int>(j)); }
    const size_t points = ij.size() / 2;
    if (count != points) throw std::invalid_argument("Reference surface-current table does not match the interface points");
    if (rows < 1) throw std::invalid_argument("Reference surface-current table has no rows");
    for (size_t t = 0; t < rows * points; ++t)
      if (!std::isfinite(table[t])) throw std::runtime_error("Nonfinite equivalent boundary current");
    if (!std::isfinite(translation_beta)) throw std::invalid_argument("Nonfinite translation beta");
    C.refTable.resize(rows * points);
    C.refPoints.resize(ij.size());
    hipCheck(hipMemcpy(C.refTable.get(), table, rows * points * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    hipCheck(hipMemcpy(C.refPoints.get(), ij.data(), ij.size() * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
    C.refRows = rows;
    C.refCount = points;
    C.refBeta = translation_beta;
    C.refCurrent = true;
  });
}

aprl_status aprl_window_complete_inlet(aprl_fields* fields, aprl_coupling* coupling, aprl_particles* particles,
                                       const double* origin) {
  return guarded([&] {
    if (!fields || !coupling || !particles || !origin) throw std::invalid_argument("Null inlet argument");
    if (coupling->fields != fields || particles->fields != fields)
      throw std::invalid_argument("Inlet, coupling and particles belong to different field grids");
    const int shape = coupling->c.shape;
    const bool xBinomial = coupling->c.smoothX;
    if (fields->fp32) completeInletSplit(fields->s32, particles->sp, origin, shape, xBinomial);
    else completeInletPlain(fields->s, particles->p, origin, shape, xBinomial);
    // the shift's inserted-plane energy is measured on the completed plane
    auto inserted = [](auto& s) {
      if (!s.ledger.insertPending) return;
      ledgerAfterShift(s);
      s.ledger.insertPending = false;
    };
    if (fields->fp32) inserted(fields->s32);
    else inserted(fields->s);
  });
}

aprl_status aprl_coupling_boundary_counters(aprl_coupling* coupling, uint64_t* counts, double* charges) {
  return guarded([&] {
    if (!coupling || !counts || !charges) throw std::invalid_argument("Null coupling or counter array");
    CouplingState& C = coupling->c;
    hipStream_t st = coupling->fields->fp32 ? coupling->fields->s32.stream : coupling->fields->s.stream;
    reserveCouplingWords(C, st);
    hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
    unsigned long long n[kCounterSlots];
    hipCheck(hipMemcpy(n, C.counterN.get(), sizeof(n), hipMemcpyDeviceToHost), "hipMemcpy");
    hipCheck(hipMemcpy(charges, C.counterQ.get(), kCounterSlots * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    for (int k = 0; k < kCounterSlots; ++k) counts[k] = n[k];
  });
}

aprl_status aprl_coupling_create(aprl_fields* fields, aprl_shape shape, int smoothX, aprl_coupling** out) {
  return guarded([&] {
    if (!fields || !out) throw std::invalid_argument("Null field grid or output handle");
    if (shape != APRL_SHAPE_LINEAR_BINOMIAL3 && shape != APRL_SHAPE_QUINTIC_COMPENSATED)
      throw std::invalid_argument("This build implements the quintic-compensated and linear-binomial3 source shapes");
    if (shape == APRL_SHAPE_LINEAR_BINOMIAL3 && !smoothX)
      throw std::invalid_argument("This build implements linear-binomial3 with the production x-binomial3 smoothing only");
    std::unique_ptr<aprl_coupling> c(new aprl_coupling);
    c->fields = fields;
    if (fields->fp32) c->c.fields32 = &fields->s32;
    else c->c.fields = &fields->s;
    c->c.shape = shape;
    // the x-binomial3 smoothing applies to the binomial3 shape only
    // (longitudinal_shape.h:109)
    c->c.smoothX = shape == APRL_SHAPE_LINEAR_BINOMIAL3 && smoothX != 0;
    const bool hasCpml = fields->fp32 ? fields->s32.hasCpml : fields->s.hasCpml;
    const aprl_cpml_parameters& cpml = fields->fp32 ? fields->s32.cpml : fields->s.cpml;
    if (hasCpml)
      for (int a = 0; a < 3; ++a) c->c.cpmlCells[a] = cpml.cells[a];
    *out = c.release();
  });
}

void aprl_coupling_destroy(aprl_coupling* coupling) {
  if (!coupling) return;
  (void)hipStreamSynchronize(coupling->fields->fp32 ? coupling->fields->s32.stream : coupling->fields->s.stream);
  delete coupling;
}

aprl_status aprl_coupling_set_frame(aprl_coupling* coupling, const aprl_boost* f) {
  return guarded([&] {
    if (!coupling || !f) throw std::invalid_argument("Null coupling or frame");
    // BoostFrameTransform::setOriginsFromGamma, boostframe.cpp:42-60.
    const double g = f->gamma;
    if (!(g >= 1.0) || !std::isfinite(g)) throw std::invalid_argument("Lorentz gamma must be finite and >= 1");
    const double gb = std::sqrt(g - 1.0) * std::sqrt(g + 1.0);
    if (!std::isfinite(f->lab_time_origin) || !std::isfinite(f->lab_z_origin) || !std::isfinite(f->box_z_origin))
      throw std::invalid_argument("Lorentz origins must be finite");
    BoostDev& B = coupling->c.frame;
    B.gamma = g;
    B.beta = gb / g;
    B.gammaBeta = gb;
    B.c0 = kC;
    B.labT0 = f->lab_time_origin;
    B.labZ0 = f->lab_z_origin;
    B.boxZ0 = f->box_z_origin;
  });
}

aprl_status aprl_coupling_add_magnet(aprl_coupling* coupling, const aprl_magnet* magnet) {
  return guarded([&] {
    if (!coupling || !magnet) throw std::invalid_argument("Null coupling or magnet");
    if (coupling->c.magnets.size() >= static_cast<size_t>(kMaxMagnets))
      throw std::invalid_argument("Too many prescribed magnets for one coupling");
    coupling->c.magnets.push_back(prepareMagnet(*magnet));
  });
}

aprl_status aprl_coupling_set_particle_boundary(aprl_coupling* coupling, const size_t* cells, double exitLabZ) {
  return guarded([&] {
    if (!coupling || !cells) throw std::invalid_argument("Null coupling or CPML cells");
    for (int a = 0; a < 3; ++a) coupling->c.cpmlCells[a] = cells[a];
    coupling->c.hasExit = std::isfinite(exitLabZ);
    coupling->c.exitLabZ = exitLabZ;
  });
}

aprl_status aprl_coupling_step(aprl_coupling* coupling, aprl_particles* particles, const double* origin,
                               double time, unsigned substeps, aprl_step_report* report) {
  return guarded([&] {
    if (!coupling || !particles || !origin) throw std::invalid_argument("Null coupling, particles or origin");
    if (particles->fields != coupling->fields)
      throw std::invalid_argument("Particles and coupling belong to different field grids");
    if (coupling->fields->fp32) couplingStepSplit(coupling->c, particles->sp, origin, time, substeps, report);
    else couplingStep(coupling->c, particles->p, origin, time, substeps, report);
  });
}

aprl_status aprl_coupling_prepare_step(aprl_coupling* coupling, aprl_particles* particles, const double* origin) {
  return guarded([&] {
    if (!coupling || !particles || !origin) throw std::invalid_argument("Null coupling, particles or origin");
    if (particles->fields != coupling->fields)
      throw std::invalid_argument("Particles and coupling belong to different field grids");
    if (coupling->fields->fp32) prepareStepSplit(coupling->c, particles->sp, origin);
    else prepareStep(coupling->c, particles->p, origin);
  });
}

aprl_status aprl_coupling_set_split_trials(aprl_coupling* coupling, unsigned trials, int verify) {
  return guarded([&] {
    if (!coupling) throw std::invalid_argument("Null coupling");
    if (!coupling->fields->fp32) throw std::invalid_argument("Fixed Picard trials apply to FP32-split grids only");
    if (trials < 2 || trials > 64) throw std::invalid_argument("FP32-split Picard trials must be in [2, 64]");
    coupling->c.splitTrials = trials;
    coupling->c.splitVerify = verify != 0;
  });
}

aprl_status aprl_coupling_set_trial_batch(aprl_coupling* coupling, unsigned batch) {
  return guarded([&] {
    if (!coupling) throw std::invalid_argument("Null coupling");
    if (coupling->fields->fp32) throw std::invalid_argument("Picard trial batches apply to plain FP64 grids only");
    if (batch > static_cast<unsigned>(kMaxTrials)) throw std::invalid_argument("Picard trial batch must be in [0, 64]");
    coupling->c.picardFixed = batch;
  });
}

}  // extern "C"
