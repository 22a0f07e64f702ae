// Driver around the reference's own hot-path translation units (TEST
// INFRASTRUCTURE ONLY). oracle/Makefile (target ref) compiles this file together with
// the reference sources where they lie under the reference checkout; the
// result goes to oracle/_ref/ and is only used to pin the oracle.
//
// Entry points mirror oracle/orc_capi.cpp with the prefix ref_. The moving
// window's reference callback reads planes staged per (component, k, step),
// so both builds consume identical inlet data.
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include <mpi.h>

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

double ref_fields_energy_density(void* h, std::size_t i, std::size_t j, std::size_t k) {
  return R(h)->f->fieldEnergyDensityCell(i, j, k);
}
double ref_fields_poynting(void* h, std::size_t i, std::size_t j, std::size_t k, unsigned axis) {
  return R(h)->f->poyntingComponentCell(i, j, k, axis);
}

}  // extern "C"
