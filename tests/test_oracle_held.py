"""Oracle held (frozen) particles pinned against the reference build.

CPU-only. prepareParticle's hold-transverse release (simulation.cpp:
4919-5035), the frozen leg and release kink of processActiveSegment
(4676-4746) and the moving-window upstream admission with the shaped
support of held particles (5121-5155) are restated on both sides; the
reference build runs them on its own RelativisticParticleSI, coupling push,
depositor and depositFrozenSourceSupport, and the two agree bitwise over
whole window steps. A held particle's z is the repeated addition of
c dt beta_held until its release.
"""
import numpy as np
import pytest

import event_cases as E
import harness as H

needs_ref = pytest.mark.skipif(not H.have_ref(), reason="reference build oracle/_ref not present")


def run_held(f, c, steps, seed=12):
    rng = np.random.default_rng(seed)
    offset = E.OFFSET0
    logs = []
    for n in range(steps):
        lower, (ex, ey), shift = E.planes_for(rng, range(3)), E.planes_for(rng, range(2)), E.planes_for(rng, range(6))
        f.clear_current()
        f.window_begin(n, offset, lower)
        f.window_advance_magnetic()
        f.window_set_inlet(offset, n + 1, ex, ey)
        logs.append(c.step(E.window_origin(offset), n * E.DZ / H.C_SI, 1))
        offset -= 1
        f.window_shift(n + 1, offset, shift)
        c.finish_window(E.window_origin(offset))
    return logs


@needs_ref
@pytest.mark.parametrize("hold", [True, False])
@pytest.mark.parametrize("shape,smooth", [("binomial3", True), ("quintic", False)])
def test_held_release_and_admission_bitwise(shape, smooth, hold):
    """Both held motions: z += c dt beta with the transverse position held,
    and the reference's default ballistic pushPosition."""
    fo, co = E.held_oracle_case("orc", shape, smooth, hold=hold)
    fr, cr = E.held_oracle_case("ref", shape, smooth, hold=hold)
    lo = run_held(fo, co, 6)
    lr = run_held(fr, cr, 6)
    assert lo == lr
    for a in range(9):
        assert np.array_equal(fo.view(a), fr.view(a)), f"component {a}"
    xo, uo = co.particles()
    xr, ur = cr.particles()
    assert np.array_equal(xo, xr) and np.array_equal(uo, ur)
    assert np.array_equal(co.frozen(), cr.frozen())
    assert np.array_equal(co.ids(), cr.ids())
    fr0 = E.held_particles()[5]
    assert 0 < co.frozen().sum() < fr0.sum()  # some released, some still held


def test_ballistic_held_particles_move_along_their_momentum():
    """Without the transverse hold a held particle moves x += (c dt / gamma) u
    each step with its momentum unchanged until its release."""
    f, c = E.held_oracle_case("orc", hold=False)
    x0, u0, q, m, ids, frozen0, beta = E.held_particles()
    steps = 3
    run_held(f, c, steps)
    x, u = c.particles()
    fr = c.frozen()
    checked = 0
    for k, pid in enumerate(c.ids()):
        i = int(pid) - 1
        if not (frozen0[i] and fr[k]):
            continue
        g = np.sqrt(1.0 + np.dot(u0[i], u0[i]))
        expect = x0[i].copy()
        for _ in range(steps):
            expect = expect + (H.C_SI * (E.DZ / H.C_SI) / g) * u0[i]
        assert np.allclose(x[k], expect, rtol=0, atol=1e-15)
        assert np.array_equal(u[k], u0[i])
        checked += 1
    assert checked > 0


def test_held_particles_move_by_repeated_addition_until_release():
    f, c = E.held_oracle_case("orc")
    x0, u0, q, m, ids, frozen0, beta = E.held_particles()
    steps = 5
    run_held(f, c, steps)
    x, u = c.particles()
    fr = c.frozen()
    idn = c.ids()
    released = 0
    for k, pid in enumerate(idn):
        i = int(pid) - 1
        if not frozen0[i]:
            continue
        if fr[k]:
            z = x0[i, 2]
            inc = H.C_SI * (E.DZ / H.C_SI) * beta[i]
            for _ in range(steps):
                z = z + inc
            assert x[k, 2] == z
            assert x[k, 0] == x0[i, 0] and x[k, 1] == x0[i, 1]
            assert np.array_equal(u[k], u0[i])
        else:
            released += 1
    assert released > 0
    assert fr.sum() > 0
