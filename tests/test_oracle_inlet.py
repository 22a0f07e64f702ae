"""Oracle charge-matched inlet pinned against the reference build and the
Gauss law it enforces.

CPU-only. The reference build runs the header-only ChargeMatchedInlet
(charge_matched_inlet.h) as MovingWindowFields' new-layer completion with
the centered normal reference; the oracle restates both and must agree
bitwise over whole coupled window steps. Independently of the reference,
the completed layer satisfies eps0 (div_xy E + (Ez(1) - Ez(0)) / dz) = rho
on the aperture vertices of the inlet plane.
"""
import numpy as np
import pytest

import event_cases as E
import harness as H

needs_ref = pytest.mark.skipif(not H.have_ref(), reason="reference build oracle/_ref not present")


def run_inlet(f, c, steps, seed=6):
    rng = np.random.default_rng(seed)
    offset = E.OFFSET0
    log = []
    ez_plane = E.planes_for(rng, [2])[0]
    for n in range(steps):
        lower, (ex, ey), shift = E.planes_for(rng, range(3)), E.planes_for(rng, range(2)), E.planes_for(rng, range(6))
        f.clear_current()
        f.window_begin_centered(n, offset, lower, ez_plane)
        f.window_advance_magnetic()
        f.window_set_inlet(offset, n + 1, ex, ey)
        log.append(c.step(E.window_origin(offset), n * E.DZ / H.C_SI, 1))
        offset -= 1
        f.window_shift(n + 1, offset, shift)
        c.complete_inlet(E.window_origin(offset))
        c.finish_window(E.window_origin(offset))
        ez_plane = shift[2]  # the next begin's reference(2, offset, step)
    return log


@needs_ref
@pytest.mark.parametrize("shape,smooth", [("binomial3", True), ("quintic", False)])
def test_charge_matched_inlet_bitwise(shape, smooth):
    fo, co = E.inlet_oracle_case("orc", shape, smooth)
    fr, cr = E.inlet_oracle_case("ref", shape, smooth)
    lo = run_inlet(fo, co, 4)
    lr = run_inlet(fr, cr, 4)
    assert lo == lr
    for a in range(9):
        assert np.array_equal(fo.view(a), fr.view(a)), f"component {a}"
    xo, uo = co.particles()
    xr, ur = cr.particles()
    assert np.array_equal(xo, xr) and np.array_equal(uo, ur)


@pytest.mark.parametrize("shape,smooth", [("binomial3", True), ("quintic", False)])
def test_completed_inlet_satisfies_gauss(shape, smooth):
    """After a shift the new layer matches the shaped charge of the
    particles near the inlet: the residual of the discrete Gauss law on the
    aperture vertices is at round-off while the charge term is not."""
    f, c = E.inlet_oracle_case("orc", shape, smooth)
    run_inlet(f, c, 1)
    x, u = c.particles()
    nx, ny, nz = E.DIMS
    o = E.window_origin(E.OFFSET0 - 1)
    q = E.inlet_particles()[2][0]
    # shaped charge at the inlet plane vertices (node k = 1 of the z basis)
    rho = np.zeros((ny + 1, nx + 1))
    for p in range(len(x)):
        z = (x[p, 2] - o[2]) / E.DZ
        wz = H.shaped_node_weight(shape, z, 1)
        if wz == 0.0:
            continue
        xx, yy = (x[p, 0] - o[0]) / E.DX, (x[p, 1] - o[1]) / E.DX
        ix, j = int(np.floor(xx)), int(np.floor(yy))
        fx, fy = xx - ix, yy - j
        wx = [0.25 * (1 - fx), 0.5 - 0.25 * fx, 0.25 + 0.25 * fx, 0.25 * fx] if smooth else [1 - fx, fx]
        first = ix - 1 if smooth else ix
        for b in range(2):
            for a, w in enumerate(wx):
                rho[j + b, first + a] += q / (E.DX * E.DX * E.DZ) * w * (fy if b else 1 - fy) * wz
    ex, ey, ez = f.view(0), f.view(1), f.view(2)
    cx, cy = E.CELLS[0], E.CELLS[1]
    worst, scale = 0.0, 0.0
    for j in range(cy + 1, ny - cy):
        for i in range(cx + 1, nx - cx):
            div = ((ex[1, j, i] - ex[1, j, i - 1]) / E.DX + (ey[1, j, i] - ey[1, j - 1, i]) / E.DX
                   + (ez[1, j, i] - ez[0, j, i]) / E.DZ)
            worst = max(worst, abs(H.EPS0 * div - rho[j, i]))
            scale = max(scale, abs(rho[j, i]))
    assert scale > 0.0
    assert worst <= 1e-10 * scale, (worst, scale)
