"""GPU moving window with the pending admission queue vs the oracle.

Held electrons upstream of the inlet wait in the library's host queue and
enter the device store as held entries 8 steps before they are due (the
reach of visitNear's support deposits); the oracle runs the reference's
scheme (admission at the due step, visitNear(step, 8) support, the waiting
particles' inlet charge). FP64 to 1e-12 with identical Picard trial counts;
the particles still waiting in the oracle's queue are in the store at the
same, bitwise equal, repeated-addition z. FP32-split with its FP32 error.
"""
import numpy as np
import pytest

import event_cases as E
import harness as H

pytestmark = pytest.mark.gpu

SHAPES = {"binomial3": ("binomial3", True, "linear-binomial3"), "quintic": ("quintic", False, "quintic-compensated")}
LOOK_AHEAD = 8


def rel_l2(a, b):
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def run(shape, precision, steps=12, seed=14, inlet=False):
    import aprl_amd as A
    from aprl_amd import pending
    oshape, smooth, gshape = SHAPES[shape]
    orc, co = E.pending_oracle_case("orc", oshape, smooth)
    gpu = A.EBFieldGrid(A.EBGridGeometry(*E.DIMS, E.DX, E.DX, E.DZ, E.DZ / H.C_SI), precision=precision)
    win = A.MovingWindowFields(gpu, A.EBCPMLParameters(E.CELLS))
    if inlet:
        co.enable_charge_match(E.REF_ORIGIN)
        win.setChargeMatch(True)
    for a in range(9):
        gpu.set_component(a, orc.view(a))
    cg = A.ConservativeSourceCoupling(gpu, gshape)
    cg.setFrame(E.GAMMA_BOOST)
    cg.setParticleBoundary(E.CELLS)
    cg.setActivationPlane(E.activation_lab_z())
    queue = pending.PendingParticleQueue(pending.PendingWindowGeometry(*E.pending_geometry()))
    x, u, q, m, ids, frozen, beta = E.pending_particles()
    x, u, q, m, w, ids, frozen, beta = queue.extract(x, u, q, m, np.ones(len(q)), ids, frozen, beta)
    assert queue.size() == co.pending_state()[0]
    store = A.ParticleStore(gpu, len(q))
    store.upload(x, u, q, m, w, ids)
    store.setHeld(frozen, beta)
    rng = np.random.default_rng(seed)
    offset = E.OFFSET0
    ez_plane = E.planes_for(rng, [2])[0]
    logs = []
    for n in range(steps):
        print(f"pending window step {n}", flush=True)
        origin = E.window_origin(offset)
        lower, inlet_planes, shift = E.planes_for(rng, range(3)), E.planes_for(rng, range(2)), E.planes_for(rng, range(6))
        co.admit(n)
        queue.admit(n, store, LOOK_AHEAD)
        t = n * E.DZ / H.C_SI
        orc.clear_current(); gpu.clearCurrent()
        if inlet:
            orc.window_begin_centered(n, offset, lower, ez_plane); win.beginStepCentered(n, offset, lower, ez_plane)
        else:
            orc.window_begin(n, offset, lower); win.beginStep(n, offset, lower)
        orc.window_advance_magnetic(); gpu.advanceMagnetic()
        orc.window_set_inlet(offset, n + 1, *inlet_planes); win.setInlet(*inlet_planes)
        ro = co.step(origin, t, 1)
        rg = cg.step(store, origin, t)
        logs.append((ro[0], rg[0]))
        offset -= 1
        orc.window_shift(n + 1, offset, shift); win.shift(n + 1, offset, shift)
        if inlet:
            co.complete_inlet(E.window_origin(offset))
            win.completeInlet(cg, store, E.window_origin(offset))
        co.finish_window(E.window_origin(offset))
        ez_plane = shift[2]
    return orc, co, gpu, store, queue, logs, steps


def compare(orc, co, gpu, store, steps, tol, exact_waiting):
    for a in range(9):
        err = rel_l2(gpu.component(a), orc.view(a))
        assert err <= tol, f"component {a}: rel L2 {err:.3e}"
    xg, ug, _, _, _, idg = store.download()
    st = store.states()
    xo, uo = co.particles()
    ido = co.ids()
    fo = co.frozen()
    # the store also holds the particles the oracle's queue still keeps,
    # due within the look-ahead, at their start-of-step z
    zw, iw = co.pending_near(steps, LOOK_AHEAD - 1)
    assert len(idg) == len(ido) + len(iw)
    early = np.isin(idg, iw)
    assert np.all(st[early] == 2)
    ge, oe = np.argsort(idg[early]), np.argsort(iw)
    assert np.array_equal(idg[early][ge], iw[oe])
    if exact_waiting:
        assert np.array_equal(xg[early][ge][:, 2], zw[oe])
    else:
        assert rel_l2(xg[early][ge][:, 2], zw[oe]) <= tol
    keep = ~early
    og, oo = np.argsort(idg[keep]), np.argsort(ido)
    assert np.array_equal(idg[keep][og], ido[oo])
    assert np.array_equal((st[keep][og] == 2).astype(np.int32), fo[oo])
    ref = xo[oo] - xo.mean(0)
    assert rel_l2(xg[keep][og] - xo.mean(0), ref) <= tol
    assert rel_l2(ug[keep][og], uo[oo]) <= tol
    return fo


@pytest.mark.parametrize("shape", ["binomial3", "quintic"])
def test_pending_fp64_matches_oracle(shape):
    orc, co, gpu, store, queue, logs, steps = run(shape, "fp64")
    assert all(a == b for a, b in logs), logs
    compare(orc, co, gpu, store, steps, 1e-12, True)
    waiting, _, admitted = co.pending_state()
    assert admitted > 5 and waiting > 5
    assert queue.admittedCount() > admitted


def test_pending_with_charge_matched_inlet_fp64():
    orc, co, gpu, store, queue, logs, steps = run("binomial3", "fp64", steps=8, inlet=True)
    assert all(a == b for a, b in logs), logs
    compare(orc, co, gpu, store, steps, 1e-12, True)


def test_pending_fp32_split_tracks_oracle():
    orc, co, gpu, store, queue, logs, steps = run("binomial3", "fp32-split")
    compare(orc, co, gpu, store, steps, 2e-5, False)
