// Builds include/aprl_reference_shim.hpp against the reference's own headers
// (linked with the reference build in oracle/_ref for the constructors of its
// types) and runs what needs no GPU: the status-to-exception mapping and the
// component shapes the window planes are sized with.
#include <cstdio>

#include "aprl_reference_shim.hpp"

namespace {
int failures = 0;
void expect(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAILED: %s\n", what);
    ++failures;
  }
}
}  // namespace

int main() {
  // a null geometry is rejected with std::invalid_argument before any device work
  bool threw = false;
  try {
    aprl_fields* f = nullptr;
    aprl::gpu::check(aprl_fields_create(nullptr, APRL_FP64, &f));
  } catch (const std::invalid_argument& e) {
    threw = std::string(e.what()).size() > 0;
  }
  expect(threw, "null geometry -> std::invalid_argument with the library's message");
  // the reference's geometry, CPML parameters and boost as C-ABI structs
  aprl::EBGridGeometry g(12, 10, 16, 2e-6, 2e-6, 1e-6, 1e-6 / 299792458.0, aprl::EBMaxwellSolver::CowanZ);
  const aprl_grid_geometry c = aprl::gpu::geometryOf(g);
  expect(c.nx == 12 && c.ny == 10 && c.nz == 16 && c.solver == 1 && c.dz == 1e-6, "geometryOf");
  aprl::EBCPMLParameters p;
  p.cells[0] = 4;
  p.cells[1] = 4;
  p.cells[2] = 0;
  const aprl_cpml_parameters cp = aprl::gpu::cpmlOf(p);
  expect(cp.cells[0] == 4 && cp.polynomial_order == p.polynomialOrder && cp.kappa_max == p.kappaMax, "cpmlOf");
  aprl::BoostFrameTransform frame;
  frame.setOriginsFromGamma(5.0, 299792458.0, 1e-9, -0.1, 0.0);
  const aprl_boost b = aprl::gpu::boostOf(frame);
  expect(b.gamma == 5.0 && b.lab_time_origin == 1e-9 && b.lab_z_origin == -0.1 && b.box_z_origin == 0.0, "boostOf");
  // component shapes against the reference's own YeeComponent sizes
  aprl::EBFieldGrid host(g);
  const aprl::YeeComponent* comp[9] = {&host.ex(), &host.ey(), &host.ez(), &host.bx(), &host.by(),
                                        &host.bz(), &host.jx(), &host.jy(), &host.jz()};
  for (int a = 0; a < 9; ++a) {
    std::size_t n[3];
    aprl::gpu::componentShape(g, a, n);
    expect(n[0] == comp[a]->nx() && n[1] == comp[a]->ny() && n[2] == comp[a]->nz(), "componentShape");
  }
  // the reference's ambiguous-shape rule and the unbuilt source forms are
  // refused before any device work
  const std::size_t guards[3] = {3, 3, 2};
  const aprl::FieldVector<aprl::Double> origin;
  const std::vector<aprl::RelativisticParticleSI> none;
  const char* expected[3] = {"Ambiguous reference source shape", "The retired linear-notch reference source is not built",
                             "The open-wall transverse reference is not built"};
  const bool flags[3][4] = {{true, true, false, false}, {false, true, false, false}, {false, false, false, true}};
  for (int c = 0; c < 3; ++c) {
    std::string message;
    try {
      aprl::gpu::InletReference r(g, origin, none, 0.3, guards, flags[c][0], flags[c][1], flags[c][2], false,
                                  flags[c][3]);
    } catch (const std::invalid_argument& e) {
      message = e.what();
    }
    expect(message == expected[c], expected[c]);
  }
  std::printf(failures ? "shim check failed\n" : "shim check passed\n");
  return failures ? 1 : 0;
}
