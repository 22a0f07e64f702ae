"""Oracle particle-boundary events pinned against the reference build and
closed forms.

CPU-only. ParticleCPMLRegion::firstEntry, ParticleOpenBoundary::firstExit and
currentDampingFactor run in oracle/_ref from the reference's own
particle_boundary.cpp; the event selection of processActiveSegment /
processCarrierSegment (simulation.cpp:4310-4917) is restated on both sides
and must agree bitwise over whole coupled window steps. Closed forms pin the
hit fractions (brute force over the faces) and the damping (the three-point
Gauss rule is exact for the cubic grading, so a straight segment normal to
one CPML layer has a closed-form integral).
"""
import math

import numpy as np
import pytest

import event_cases as E
import harness as H

needs_ref = pytest.mark.skipif(not H.have_ref(), reason="reference build oracle/_ref not present")


make_case = E.oracle_case


def run_window(f, c, steps, seed=4, exit_plane=False):
    """Window steps with the origin at fma(offset, dz, reference origin) and
    the particle part of finishMovingWindowStep after every shift."""
    rng = np.random.default_rng(seed)
    offset = E.OFFSET0
    exit_z = E.exit_lab_z() if exit_plane else float("inf")
    log = []
    for n in range(steps):
        lower, (ex, ey), shift = E.planes_for(rng, range(3)), E.planes_for(rng, range(2)), E.planes_for(rng, range(6))
        f.clear_current()
        f.window_begin(n, offset, lower)
        f.window_advance_magnetic()
        f.window_set_inlet(offset, n + 1, ex, ey)
        log.append(c.step(E.window_origin(offset), n * E.DZ / H.C_SI, 1, exit_z))
        offset -= 1
        f.window_shift(n + 1, offset, shift)
        c.finish_window(E.window_origin(offset))
    return log


def state(c):
    x, u = c.particles()
    return x, u, c.ids(), c.carriers(), c.counters()


@needs_ref
@pytest.mark.parametrize("shape,smooth", [("binomial3", True), ("quintic", False)])
@pytest.mark.parametrize("exit_plane", [False, True])
def test_event_window_steps_bitwise(shape, smooth, exit_plane):
    """CPML entries, outer exits, window exits, carriers and crops over
    whole coupled window steps: oracle == reference build, bitwise."""
    fo, co = make_case("orc", shape, smooth, exit_plane)
    fr, cr = make_case("ref", shape, smooth, exit_plane)
    lo = run_window(fo, co, 6, exit_plane=exit_plane)
    lr = run_window(fr, cr, 6, exit_plane=exit_plane)
    assert lo == lr
    for a in range(9):
        assert np.array_equal(fo.view(a), fr.view(a)), f"component {a}"
    so, sr = state(co), state(cr)
    for k in range(3):
        assert np.array_equal(so[k], sr[k])
    for k in range(5):
        assert np.array_equal(so[3][k], sr[3][k]), f"carrier field {k}"
    assert np.array_equal(so[4][0], sr[4][0]) and np.array_equal(so[4][1], sr[4][1])
    counts = so[4][0]
    # every kind of event happened
    assert counts[0] > 0 and counts[3] > 0, counts           # x- and y+ CPML entries
    if exit_plane:
        assert counts[18] > 0 and counts[17] > 0, counts     # window exits, carriers leaving z+
    else:
        assert counts[11] > 0, counts                        # particle exits through z+
    assert len(so[3][0]) > 0                                 # live carriers


def test_event_bookkeeping_invariants():
    """Every particle is retained, became a carrier (counted at its CPML or
    window entry) or left through an outer face; counted charges are the
    particle charges (carrier exits and crops carry their weights <= 1)."""
    f, c = make_case("orc", exit_plane=True)
    n0 = len(c.ids())
    q0 = E.particles(exit_plane=True)[2][0]
    run_window(f, c, 6, exit_plane=True)
    counts, charges = c.counters()
    retained = len(c.ids())
    entered = int(counts[0:6].sum() + counts[18])
    direct = int(counts[6:12].sum())
    assert retained + entered + direct == n0
    assert math.isclose(charges[0:6].sum() + charges[18], entered * q0, rel_tol=1e-14)
    assert math.isclose(charges[6:12].sum(), direct * q0, rel_tol=1e-14)
    xc, vc, qc, wc, idc = c.carriers()
    assert len(idc) + int(counts[12:18].sum() + counts[19]) <= entered
    assert np.all((wc > 0) & (wc <= 1))
    # carriers move on straight lines at |v| < c
    assert np.all(np.linalg.norm(vc, axis=1) < H.C_SI)
    # carriers in the transverse CPML are damped
    assert np.any(wc < 1.0)
    ids = c.ids()
    assert not set(ids.tolist()) & set(idc.tolist())


@needs_ref
def test_boundary_queries_bitwise_and_closed_form():
    """firstEntry / firstExit / currentDampingFactor: oracle == reference on
    random segments; hit fractions equal the brute-force first crossing; the
    damping of a segment normal to the x- layer has its closed form."""
    fo, co = make_case("orc")
    fr, cr = make_case("ref")
    o = E.window_origin(E.OFFSET0)
    nx, ny, nz = E.DIMS
    ext = np.array([nx * E.DX, ny * E.DX, nz * E.DZ])
    rng = np.random.default_rng(2)
    phys_lo = o + np.array([E.CELLS[0] * E.DX, E.CELLS[1] * E.DX, 0.0])
    phys_hi = o + ext - np.array([E.CELLS[0] * E.DX, E.CELLS[1] * E.DX, 0.0])
    hits = 0
    for t in range(400):
        a = phys_lo + rng.uniform(0.0, 1.0, 3) * (phys_hi - phys_lo)
        b = a + rng.uniform(-1.0, 1.0, 3) * np.array([6 * E.DX, 6 * E.DX, 6 * E.DZ])
        for kind in (0, 1):
            ro = co.boundary_query(o, kind, a, b)
            rr = cr.boundary_query(o, kind, a, b)
            assert np.array_equal(ro, rr), (kind, a, b)
            lo_, hi_ = (phys_lo, phys_hi) if kind == 0 else (o, o + ext)
            active = [0, 1] if kind == 0 else [0, 1, 2]
            best, face = math.inf, -1
            for ax in active:
                if b[ax] < lo_[ax]:
                    f_ = (lo_[ax] - a[ax]) / (b[ax] - a[ax])
                    if f_ < best:
                        best, face = f_, 2 * ax
                elif b[ax] > hi_[ax]:
                    f_ = (hi_[ax] - a[ax]) / (b[ax] - a[ax])
                    if f_ < best:
                        best, face = f_, 2 * ax + 1
            assert bool(ro[0]) == (face >= 0)
            if face >= 0:
                hits += 1
                # the reference's faces are origin + n h (CPML) and fma(n, h, origin)
                assert math.isclose(ro[1], min(1.0, max(0.0, best)), rel_tol=1e-12, abs_tol=1e-15)
                assert int(ro[2]) == face
        dur = rng.uniform(0, 2) * E.DZ / H.C_SI
        ro = co.boundary_query(o, 2, a, b, dur)
        rr = cr.boundary_query(o, 2, a, b, dur)
        assert ro[0] == rr[0]
    assert hits > 100
    # closed form: x from depth d0 to d1 inside the x- layer, cubic grading
    order, R = 3.0, 1e-8
    thickness = E.CELLS[0] * E.DX
    rate_max = -(order + 1.0) * H.C_SI * math.log(R) / (2.0 * thickness)
    y, z = o[1] + 0.5 * ext[1], o[2] + 0.5 * ext[2]
    for d0, d1 in [(0.0, 0.5), (0.2, 0.9), (0.6, 0.1)]:
        a = np.array([phys_lo[0] - d0 * thickness, y, z])
        b = np.array([phys_lo[0] - d1 * thickness, y, z])
        dur = 0.7 * E.DZ / H.C_SI
        mean = rate_max * (d1 ** 4 - d0 ** 4) / (4.0 * (d1 - d0))
        assert math.isclose(co.boundary_query(o, 2, a, b, dur)[0], math.exp(-dur * mean), rel_tol=1e-13)
    # no damping inside the physical region
    a = phys_lo + 0.3 * (phys_hi - phys_lo)
    assert co.boundary_query(o, 2, a, a + E.DZ, 1e-15)[0] == 1.0
