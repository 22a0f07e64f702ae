// Oracle C-ABI for the test harness (TEST INFRASTRUCTURE ONLY; never linked
// into the product). Entry points mirror oracle/ref_driver.cpp (prefix ref_)
// so the same Python test drives the restatement and the reference build.
#include <cstring>
#include <string>

#include "orc_fields.hpp"
#include "orc_particles.hpp"

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

double orc_fields_energy_density(void* h, std::size_t i, std::size_t j, std::size_t k) {
  return F(h)->energyDensityCell(i, j, k);
}
double orc_fields_poynting(void* h, std::size_t i, std::size_t j, std::size_t k, unsigned axis) {
  return F(h)->poyntingCell(i, j, k, axis);
}

}  // extern "C"
