"""The moving window's pending admission queue: repeatHeldAddition and
PendingParticleQueue pinned three ways.

CPU-only. The oracle restates repeatHeldAddition by its definition (count
literal binary64 additions) and the queue with a step-by-step due scan; the
reference build runs pending_particle_queue.cpp itself (oracle/_ref); the
library's host queue (aprl_pending_*, pure host code, no GPU call) jumps
through whole binades. All three agree bitwise. Over whole coupled window
steps the oracle's queue (admission before the step, visitNear(step, 8)
support deposits, the inlet charge of the waiting particles) equals the
reference build bitwise, and equals resident held particles (no queue) to
round-off, which is what the library's early admission relies on.
"""
import math

import numpy as np
import pytest

import event_cases as E
import harness as H

needs_ref = pytest.mark.skipif(not H.have_ref(), reason="reference build oracle/_ref not present")


def product():
    from aprl_amd import pending
    return pending


def repeat_cases(seed=3, count=600):
    """(initial, increment, count) covering ties, binade crossings in both
    directions, zero crossings, negative increments and held-window scales."""
    rng = np.random.default_rng(seed)
    out = []
    for t in range(count):
        kind = t % 8
        if kind == 0:    # held particle upstream of the window
            x, inc = -rng.uniform(1e-6, 4e-5), 0.99e-6 * rng.uniform(0.5, 1.0)
        elif kind == 1:  # exact half-ulp ties
            x = rng.uniform(-1, 1) * 2.0 ** int(rng.integers(-40, 40))
            inc = (int(rng.integers(0, 12)) + 0.5) * np.spacing(abs(x))
        elif kind == 2:  # start on a negative power of two
            x = -(2.0 ** int(rng.integers(-30, 10)))
            inc = np.spacing(abs(x)) * rng.uniform(0.2, 9.0)
        elif kind == 3:  # negative increments
            x, inc = rng.uniform(-10, 10), -rng.uniform(1e-9, 1e-2)
        elif kind == 4:  # increments below half an ulp: a fixed point
            x = rng.uniform(1, 2) * 2.0 ** int(rng.integers(-5, 5))
            inc = np.spacing(x) * rng.uniform(0.0, 0.5)
        elif kind == 5:  # through zero
            x, inc = -rng.uniform(1e-9, 1e-6), rng.uniform(1e-12, 1e-9)
        elif kind == 6:  # dyadic values: sums are exact
            x, inc = float(rng.integers(-1000, 1000)) * 2.0 ** -20, float(rng.integers(1, 64)) * 2.0 ** -26
        else:            # large magnitudes
            x, inc = rng.uniform(-1, 1) * 1e200, rng.uniform(1e180, 1e186)
        n = int(rng.integers(0, 4000))
        out.append((float(x), float(inc), n))
    out += [(-0.0, 0.0, 3), (-0.0, 0.0, 0), (1.0, -1.0, 1), (-1.0, 1.0, 1), (0.0, 5e-324, 7),
            (-5e-324, 5e-324, 2), (1.0, 2.0 ** -53, 1000), (1.0 - 2.0 ** -53, 2.0 ** -54, 10)]
    return out


def same(a, b):
    return a == b and math.copysign(1.0, a) == math.copysign(1.0, b)


def test_repeat_held_addition_library_equals_literal_additions():
    P = product()
    for x, inc, n in repeat_cases():
        lit = H.repeat_held_addition("orc", x, inc, n)
        assert same(P.repeatHeldAddition(x, inc, n), lit), (x, inc, n)


def test_repeat_held_addition_long_runs_and_closed_forms():
    P = product()
    # millions of additions against the literal loop
    for x, inc, n in [(-4.0e-5, 0.99e-6, 3_000_000), (1.0, 1.0e-16 * 3, 2_000_000), (-3.7, 1.1e-15, 4_000_000),
                      (2.0 ** 30, 0.75, 5_000_000)]:
        assert same(P.repeatHeldAddition(x, inc, n), H.repeat_held_addition("orc", x, inc, n)), (x, inc, n)
    # dyadic sums are exact: x + n inc
    for x, inc, n in [(0.0, 2.0 ** -10, 12345), (-1.0, 2.0 ** -20, 3 * 2 ** 20), (3.0, -2.0 ** -8, 999)]:
        assert P.repeatHeldAddition(x, inc, n) == x + n * inc
    # below half an ulp nothing moves; exactly half an ulp from an even
    # significand is a tie that rounds back to it
    assert P.repeatHeldAddition(1.0, 2.0 ** -54, 10 ** 12) == 1.0
    assert P.repeatHeldAddition(1.0, 2.0 ** -53, 10 ** 12) == 1.0
    # from an odd significand the first tie goes up to even, then stays
    odd = 1.0 + 2.0 ** -52
    assert P.repeatHeldAddition(odd, 2.0 ** -53, 10 ** 12) == 1.0 + 2.0 ** -51
    # 1.5 ulp: every addition rounds to 2 ulp (ties to even from an even start)
    assert P.repeatHeldAddition(1.0, 1.5 * 2.0 ** -52, 1000) == 1.0 + 2000 * 2.0 ** -52


def test_repeat_held_addition_errors():
    P = product()
    from aprl_amd import _capi
    with pytest.raises(_capi.InvalidArgument):
        P.repeatHeldAddition(float("nan"), 1.0, 1)
    with pytest.raises(_capi.InvalidArgument):
        P.repeatHeldAddition(0.0, float("inf"), 1)
    with pytest.raises(_capi.AprlError, match="overflow"):
        P.repeatHeldAddition(1.7e308, 1e307, 10)
    assert P.repeatHeldAddition(float("-0.0"), 1.0, 0) == 0.0


@needs_ref
def test_repeat_held_addition_reference_build():
    P = product()
    for x, inc, n in repeat_cases(seed=9, count=400):
        r = H.repeat_held_addition("ref", x, inc, n)
        assert same(P.repeatHeldAddition(x, inc, n), r), (x, inc, n)
        assert same(H.repeat_held_addition("orc", x, inc, n), r), (x, inc, n)
    for x, inc, n in [(-4.0e-5, 0.99e-6, 10 ** 9), (-3.7, 1.1e-15, 10 ** 10), (1.0, 3e-16, 10 ** 12)]:
        assert same(P.repeatHeldAddition(x, inc, n), H.repeat_held_addition("ref", x, inc, n)), (x, inc, n)


def library_queue():
    P = product()
    g = P.PendingWindowGeometry(*E.pending_geometry())
    q = P.PendingParticleQueue(g)
    x, u, qq, m, ids, frozen, beta = E.pending_particles()
    kept = q.extract(x, u, qq, m, np.ones(len(qq)), ids, frozen, beta)
    return q, kept


def test_library_queue_extracts_waiting_held_particles():
    q, kept = library_queue()
    x, u, qq, m, ids, frozen, beta = E.pending_particles()
    inlet = q.origin(0)
    assert math.isclose(inlet, E.window_origin(E.OFFSET0)[2], abs_tol=1e-20)
    c_dt = E.pending_geometry()[3]
    waiting = [i for i in range(len(ids)) if frozen[i] and x[i, 2] < inlet and beta[i] >= 0
               and not (x[i, 2] + c_dt * beta[i] >= inlet)]
    assert q.size() == len(waiting) > 10
    kept_ids = kept[5]
    assert sorted(set(kept_ids.tolist()) | {int(ids[i]) for i in waiting}) == ids.tolist()
    # the rest keep their order and values
    order = [int(i) for i in ids if int(i) not in {int(ids[w]) for w in waiting}]
    assert kept_ids.tolist() == order
    # due steps: a particle is visited by view(s, 0) from its first step s
    # whose held end reaches the inlet of s
    for s in range(0, 20):
        xv, vid = q.visitNear(s, 0)
        for zi, pid in zip(xv[:, 2], vid):
            i = int(pid) - 1
            assert zi == H.repeat_held_addition("orc", x[i, 2], c_dt * beta[i], s)
            assert zi + c_dt * beta[i] >= q.origin(s)
    with pytest.raises(Exception, match="only be initialized once"):
        q.extract(x, u, qq, m, np.ones(len(qq)), ids, frozen, beta)


@needs_ref
def test_library_queue_matches_reference_queue():
    """Due steps, admission order and admitted z: the library's view over
    the reference build's queue (admitting step by step)."""
    q, _ = library_queue()
    _, cr = E.pending_oracle_case("ref")
    _, co = E.pending_oracle_case("orc")
    assert q.size() == cr.pending_state()[0] == co.pending_state()[0]
    assert q.nextDueStep() == cr.pending_state()[1] == co.pending_state()[1]
    admitted = set()
    for s in range(0, 30):
        for L in (0, 3, 8):
            zr, ir = cr.pending_near(s, L)
            zo, io = co.pending_near(s, L)
            xl, il = q.visitNear(s, L)
            zl = xl[:, 2]
            keep = np.array([int(i) not in admitted for i in il], dtype=bool)
            assert np.array_equal(ir, il[keep]) and np.array_equal(io, ir), (s, L)
            assert np.array_equal(zr, zl[keep]) and np.array_equal(zo, zr), (s, L)
        nr, no = cr.admit(s), co.admit(s)
        assert nr == no
        assert len(cr.pending_near(s, 0)[1]) == 0  # nothing due by s still waits
        ids_r, ids_o = cr.ids(), co.ids()
        assert np.array_equal(ids_r, ids_o)
        assert np.array_equal(cr.particles()[0], co.particles()[0])
        admitted |= set(int(i) for i in ids_r[len(ids_r) - nr:])
        assert cr.pending_state() == co.pending_state()
    assert cr.pending_state()[2] > 10


def run_pending(f, c, steps, seed=14, queue=True, inlet=False):
    rng = np.random.default_rng(seed)
    offset = E.OFFSET0
    logs = []
    ez_plane = E.planes_for(rng, [2])[0]
    for n in range(steps):
        lower, (ex, ey), shift = E.planes_for(rng, range(3)), E.planes_for(rng, range(2)), E.planes_for(rng, range(6))
        if queue:
            c.admit(n)
        f.clear_current()
        if inlet:
            f.window_begin_centered(n, offset, lower, ez_plane)
        else:
            f.window_begin(n, offset, lower)
        f.window_advance_magnetic()
        f.window_set_inlet(offset, n + 1, ex, ey)
        logs.append(c.step(E.window_origin(offset), n * E.DZ / H.C_SI, 1))
        offset -= 1
        f.window_shift(n + 1, offset, shift)
        if inlet:
            c.complete_inlet(E.window_origin(offset))
        c.finish_window(E.window_origin(offset))
        ez_plane = shift[2]
    return logs


@needs_ref
@pytest.mark.parametrize("shape,smooth,inlet", [("binomial3", True, False), ("quintic", False, False),
                                                ("binomial3", True, True)])
def test_pending_window_steps_bitwise(shape, smooth, inlet):
    fo, co = E.pending_oracle_case("orc", shape, smooth)
    fr, cr = E.pending_oracle_case("ref", shape, smooth)
    if inlet:
        co.enable_charge_match(E.REF_ORIGIN)
        cr.enable_charge_match(E.REF_ORIGIN)
    lo = run_pending(fo, co, 12, inlet=inlet)
    lr = run_pending(fr, cr, 12, inlet=inlet)
    assert lo == lr
    for a in range(9):
        assert np.array_equal(fo.view(a), fr.view(a)), f"component {a}"
    xo, uo = co.particles()
    xr, ur = cr.particles()
    assert np.array_equal(xo, xr) and np.array_equal(uo, ur)
    assert np.array_equal(co.ids(), cr.ids())
    assert np.array_equal(co.frozen(), cr.frozen())
    assert co.pending_state() == cr.pending_state()
    waiting, _, admitted = co.pending_state()
    assert admitted > 5 and waiting > 5


def test_pending_queue_equals_resident_held_particles():
    """The queue (admission at the due step, visitNear(step, 8) support)
    and resident held particles (their own upstream support) give the same
    step to round-off: the waiting particles' support is the same current
    in another summation order."""
    fq, cq = E.pending_oracle_case("orc", queue=True)
    fs, cs = E.pending_oracle_case("orc", queue=False)
    run_pending(fq, cq, 10)
    run_pending(fs, cs, 10, queue=False)
    for a in range(9):
        ref = fs.view(a)
        err = np.linalg.norm(fq.view(a) - ref) / max(np.linalg.norm(ref), 1e-300)
        assert err <= 1e-13, f"component {a}: {err:.3e}"
    xq, uq = cq.particles()
    xs, us = cs.particles()
    iq, is_ = cq.ids(), cs.ids()
    # the resident list also holds the particles still waiting in the queue
    waiting = cq.pending_near(10, E.PENDING_MAXIMUM_STEP)[1]
    assert len(is_) == len(iq) + len(waiting)
    common = np.isin(is_, iq)
    oq, os_ = np.argsort(iq), np.argsort(is_[common])
    assert np.array_equal(iq[oq], is_[common][os_])
    xr, ur = xs[common][os_], us[common][os_]
    mean = xr.mean(0)
    assert np.linalg.norm(xq[oq] - xr) / np.linalg.norm(xr - mean) <= 1e-12
    assert np.linalg.norm(uq[oq] - ur) / np.linalg.norm(ur) <= 1e-12
    assert np.array_equal(cq.frozen()[oq], cs.frozen()[common][os_])
