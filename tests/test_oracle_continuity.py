"""Charge continuity of the oracle's coupled step with the shaped charge.

The current of one coupled window step (pinned bitwise to the reference
build in test_oracle_particles.py) must satisfy
(rho1 - rho0)/dt + div J = 0 at every vertex to round-off, where rho is the
charge of the same source shape at the step's start and end positions:
binomial3-x -- x nodeWeights (4) x linear y (2) x binomial3 z basis (4)
(verification/transverse_binomial_checks.cpp:33-40); quintic -- linear x,
y x the compensated quintic z basis (12). This pins the charge shape that
the GPU continuity and Gauss diagnostics deposit.
"""
import numpy as np
import pytest

import harness as H
from aprl_amd import bunch

DZ = 1.0e-6


@pytest.mark.parametrize("shape,smooth", [("binomial3", True), ("quintic", False)])
def test_coupled_step_conserves_shaped_charge(shape, smooth):
    dims, cells = (16, 16, 40), (4, 4, 0)
    dx = 2.5 * DZ
    h = (dx, dx, DZ)
    dt = DZ / H.C_SI
    orc = H.CpuFields("orc", *dims, dx, dx, DZ, dt, 1)
    orc.install_window(cells)
    H.seeded_fields(orc, 3, scale=1e5)
    co = H.CpuCoupling(orc, shape, smooth, cpml_cells=cells)
    co.set_frame(5.0)
    co.add_magnet("dipole", [0, 0, -1.0], 2.0, peak=0.5)
    x, u, q, m, ids = bunch.seeded_gaussian(60, 11, [8 * dx, 8 * dx, 18 * DZ], [dx, dx, 3 * DZ], [0.0, 0.0, 10.0],
                                            [0.02, 0.02, 0.05], 3.0e5)
    co.set_particles(x, u, q, m, np.ones(60), ids)
    zero = [np.zeros(int(np.prod(H.component_shape(*dims, a)[1:]))) for a in range(6)]
    orc.clear_current()
    orc.window_begin(0, 50, zero[:3])
    orc.window_advance_magnetic()
    orc.window_set_inlet(50, 1, zero[0], zero[1])
    x0, _ = co.particles()
    co.step(np.zeros(3), 0.0)
    x1, _ = co.particles()
    rho0 = H.shaped_charge(shape, x0, q, dims, h)
    rho1 = H.shaped_charge(shape, x1, q, dims, h)
    residual, rate = H.continuity(rho0, rho1, orc.view(6), orc.view(7), orc.view(8), dt, h)
    assert rate > 0
    assert residual <= 1e-12 * rate, (residual, rate)
    # a plausible slip -- the charge at the step end taken one cell upstream
    # in z -- is caught
    rho_bad = np.roll(rho1, 1, axis=0)
    bad, _ = H.continuity(rho0, rho_bad, orc.view(6), orc.view(7), orc.view(8), dt, h)
    assert bad > 1e-3 * rate
