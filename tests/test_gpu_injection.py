"""The laboratory-plane input route on the GPU (aprl_bunch_*) against the
oracle's restatement (oracle/orc_bunch.hpp, bitwise equal to the reference
build's SIBunchPreprocessor, tests/test_oracle_bunch.py):

* the device plane-arrival generator against the same splitmix64 draws in
  numpy (the device's log and cos against numpy's: 1e-14);
* projection (with and without the transverse hold), placement, the
  frozen-injection frame origin, the boost to box time zero and the frozen
  flags on identical inputs: coordinates, momenta and held speeds within
  4 ulp of the oracle's (the three long-double sums of the projection are
  double-double sums here; the device hypot in the held speed), reports
  within 1e-15 relative; the reference's error messages;
* a production window run: the events of a seeded Gaussian bunch through
  the device input route, the held snapshot into the library's pending
  queue and the store, then 40 moving-window steps in which the queue
  admits the particles, they release at the activation plane and leave
  through the lab exit plane, against the oracle stepping the same
  snapshot with the reference's queue: E, B and J (as vector fields) and
  the particles within 1e-12 relative L2 of the oracle after every 8 steps, Picard trial counts and
  window-exit counters equal.
"""
import numpy as np
import pytest

import bunch_route as B
import config_cases as CC
import harness as H

pytestmark = pytest.mark.gpu

C_SI = H.C_SI
GAMMA_B, DZ, DX = 5.0, 1.0e-6, 2.5e-6
INPUT = -0.01
CENTER = INPUT + 60e-6


def events(n=256, seed=5, sigma_t=1e-6 / C_SI):
    return B.plane_events(n, seed, (4e-6, 4e-6), sigma_t, 50.0, 1e-3, 2e-8, 2e5)


def device_bunch(x, u, q, m, w, ids):
    import aprl_amd as A
    b = A.LabPlaneBunch(len(q))
    b.upload(x, u, q, m, w, ids)
    return b


def close_ulps(a, b, ulps=4):
    a, b = np.asarray(a), np.asarray(b)
    return np.all(np.abs(a - b) <= ulps * np.spacing(np.maximum(np.abs(a), np.abs(b))))


def test_plane_event_generator_matches_numpy_draws():
    import aprl_amd as A
    n = 5000
    x, u, q, m, w, ids = events(n, seed=77)
    b = A.LabPlaneBunch()
    b.generatePlaneEvents(n, (0.0, 0.0), (4e-6, 4e-6), 1e-6 / C_SI, (0.0, 0.0, np.sqrt(50.0 ** 2 - 1.0)),
                          (2e-8 / 4e-6, 2e-8 / 4e-6, 1e-3 * 50.0), 2e5, 77)
    xg, ug, qg, mg, wg, idg, bg = b.download()
    assert np.array_equal(idg, ids) and np.array_equal(qg, q) and np.array_equal(mg, m)
    assert np.all(wg == 1.0) and np.all(bg == 0.0)
    for a in range(3):
        assert np.allclose(xg[:, a], x[:, a], rtol=1e-13, atol=1e-13 * np.abs(x[:, a]).max())
        assert np.allclose(ug[:, a], u[:, a], rtol=1e-13, atol=1e-13 * np.abs(u[:, a]).max())


@pytest.mark.parametrize("hold", [True, False])
def test_input_route_stages_match_oracle(hold):
    x, u, q, m, w, ids = events(3000, seed=9, sigma_t=3e-6 / C_SI)
    w = 0.5 + np.arange(len(w)) % 5 * 0.25
    print(f"input route stages hold={hold}", flush=True)
    b = device_bunch(x, u, q, m, w, ids)
    rg = b.projectLabPlaneEventsToSnapshot(INPUT, CENTER, 2e-6, hold)
    xo = np.array(x)
    ro = B.project("orc", xo, u, w, INPUT, CENTER, 2e-6, hold)
    xg = b.download()[0]
    keys = ["input_plane_lab", "reference_position_lab", "minimum_forward_distance", "maximum_forward_distance",
            "recommended_minimum_reference_position", "reference_time_offset_lab", "mean_longitudinal_beta"]
    assert np.allclose([rg[k] for k in keys], ro[:7], rtol=1e-15, atol=0), (rg, ro)
    assert rg["particles"] == ro[7]
    assert close_ulps(xg, xo), np.abs(xg - xo).max()
    # placement, frame origin, boost and freeze from the same snapshot
    b.upload(xo, u, q, m, w, ids)
    rl = b.placeRelativeLabSnapshot(CENTER, -0.009, GAMMA_B * DZ)
    rlo = B.place("orc", xo, -0.009, CENTER, GAMMA_B * DZ)
    assert [rl[k] for k in ["relative_tail_lab", "relative_head_lab", "reference_position_lab", "head_after_lab",
                            "actual_head_distance", "recommended_maximum_reference_position"]] == list(rlo[:6])
    assert np.array_equal(b.download()[0], xo)
    frame = b.frozenFrameOrigin(GAMMA_B, INPUT, INPUT - 1e-6, DZ, hold)
    fo = B.frozen_origin("orc", xo, u, GAMMA_B, INPUT, INPUT - 1e-6, DZ, hold)
    assert close_ulps(frame[1:3], fo) and frame[0] == GAMMA_B and frame[3] == 0.0
    frame4 = np.array([GAMMA_B, fo[0], fo[1], 0.0])
    rb = b.boostLabSnapshotToBoxTimeZero(frame4, 0.0, hold, INPUT)
    ub_o = np.array(u)
    rbo = B.boost("orc", xo, ub_o, frame4, 0.0, hold, INPUT)
    xg, ug = b.download()[:2]
    assert close_ulps(xg, xo), np.abs(xg - xo).max()
    assert close_ulps(ug, ub_o), np.abs(ug - ub_o).max()
    bkeys = ["earliest_lab_event_time", "latest_lab_event_time", "maximum_absolute_drift_time", "minimum_lab_gamma",
             "maximum_lab_gamma"]
    assert close_ulps([rb[k] for k in bkeys], rbo[:5])
    assert rb["maximum_relative_momentum_round_trip_error"] <= 1e-15
    b.upload(xo, ub_o, q, m, w, ids)
    b.freeze(frame4, INPUT - 1e-6, hold)
    beta_o = B.freeze("orc", xo, ub_o, frame4, INPUT - 1e-6, hold)
    assert close_ulps(b.download()[6], beta_o)


def test_input_route_errors_match_oracle():
    import aprl_amd as A
    x, u, q, m, w, ids = events(200)
    b = device_bunch(x, u, q, m, w, ids)
    with pytest.raises(A.AprlError) as e:
        b.projectLabPlaneEventsToSnapshot(INPUT, INPUT - 1e-3, 0.0, True)
    with pytest.raises(H.CpuError) as eo:
        B.project("orc", np.array(x), u, w, INPUT, INPUT - 1e-3, 0.0, True)
    assert str(eo.value) in str(e.value)
    bad = w.copy()
    bad[3] = -1.0
    b.upload(x, u, q, m, bad, ids)
    with pytest.raises(A.AprlError, match="finite positive weights"):
        b.projectLabPlaneEventsToSnapshot(INPUT, CENTER, 0.0, True)
    # a snapshot already at the activation plane at box time zero
    b.upload(np.zeros((len(q), 3)), u, q, m, w, ids)
    with pytest.raises(A.AprlError, match="upstream of the input plane at box time zero"):
        b.freeze((GAMMA_B, 0.0, 0.0, 0.0), 0.0, True)


# ---------------------------------------------------------------- production run
DIMS, CELLS = (24, 24, 64), (4, 4, 0)
OFFSET0 = 50
REF_ORIGIN_Z = -(OFFSET0 + 10) * DZ  # the window spans box z [-10, 54) dz at step 0
STEPS = 40


def window_origin(offset):
    import event_cases as E
    return np.array([-0.5 * DIMS[0] * DX, -0.5 * DIMS[1] * DX, E._fma(float(offset), DZ, REF_ORIGIN_Z)])


def production_snapshot():
    """The device input route on the events of a seeded Gaussian bunch;
    returns the held box-frame snapshot and the frame."""
    import aprl_amd as A
    b = A.LabPlaneBunch()
    n = 256
    b.generatePlaneEvents(n, (0.0, 0.0), (4e-6, 4e-6), 1e-6 / C_SI, (0.0, 0.0, np.sqrt(50.0 ** 2 - 1.0)),
                          (2e-8 / 4e-6, 2e-8 / 4e-6, 0.05), 2e5, 5)
    planes = A.InjectionPlanes(input_plane_z=INPUT, activation_plane_z=INPUT, initial_center_z=CENTER)
    frame, reports = A.frozenLabPlaneInjection(b, planes, GAMMA_B, DZ)
    x, u, q, m, w, ids, beta = b.download()
    return x, u, q, m, w, ids, beta, frame


def test_production_window_run_matches_oracle():
    import aprl_amd as A
    from aprl_amd import pending
    x, u, q, m, w, ids, beta, frame = production_snapshot()
    print("production snapshot: box z cells", x[:, 2].min() / DZ, x[:, 2].max() / DZ, flush=True)
    assert np.all(beta > 0.9) and x[:, 2].max() < 0.0
    exit_lab = frame[2] + GAMMA_B * 40 * DZ  # box z 40 dz at box time zero
    frozen = np.ones(len(q), np.int32)
    # oracle: the reference's queue inside the coupling
    orc = H.CpuFields("orc", *DIMS, DX, DX, DZ, DZ / C_SI, 1)
    orc.install_window(CELLS)
    co = H.CpuCoupling(orc, "binomial3", True, cpml_cells=CELLS)
    co.set_frame(*frame)
    co.set_particles(x, u, q, m, w, ids)
    co.set_held(frozen, beta)
    co.set_activation(INPUT)
    co.set_hold_transverse(True)
    geometry = (REF_ORIGIN_Z, OFFSET0, DZ, C_SI * (DZ / C_SI), STEPS + 20)
    co.set_pending(*geometry)
    # GPU: the library's queue in front of the store
    gpu = A.EBFieldGrid(A.EBGridGeometry(*DIMS, DX, DX, DZ, DZ / C_SI))
    win = A.MovingWindowFields(gpu, A.EBCPMLParameters(CELLS))
    cg = A.ConservativeSourceCoupling(gpu, "linear-binomial3")
    cg.setFrame(*frame)
    cg.setParticleBoundary(CELLS, exit_lab)
    cg.setActivationPlane(INPUT)
    cg.setHoldTransverse(True)
    queue = pending.PendingParticleQueue(pending.PendingWindowGeometry(*geometry))
    xs, us, qs, ms, ws, idss, fs, bs = queue.extract(x, u, q, m, w, ids, frozen, beta)
    queued = queue.size()
    assert queued == co.pending_state()[0] and queued > 50
    store = A.ParticleStore(gpu, len(q))
    store.upload(xs, us, qs, ms, ws, idss)
    store.setHeld(fs, bs)
    zero = {a: np.zeros(int(np.prod(H.component_shape(*DIMS, a)[1:]))) for a in range(6)}
    lower, inlet, shift = [zero[a] for a in range(3)], [zero[a] for a in range(2)], [zero[a] for a in range(6)]
    offset = OFFSET0
    trials = []
    for n in range(STEPS):
        origin = window_origin(offset)
        if n:
            # the device crops the shifted window's carriers when the next
            # step starts, at the origin finishMovingWindowStep uses
            co.finish_window(origin)
        co.admit(n)
        queue.admit(n, store, 8)
        t = n * DZ / C_SI
        orc.clear_current(); gpu.clearCurrent()
        orc.window_begin(n, offset, lower); win.beginStep(n, offset, lower)
        orc.window_advance_magnetic(); gpu.advanceMagnetic()
        orc.window_set_inlet(offset, n + 1, *inlet); win.setInlet(*inlet)
        ro = co.step(origin, t, 1, exit_lab)
        rg = cg.step(store, origin, t)
        trials.append((ro[0], rg[0]))
        offset -= 1
        orc.window_shift(n + 1, offset, shift); win.shift(n + 1, offset, shift)
        if (n + 1) % 8 == 0:
            errs = CC.vector_errors(gpu, orc)
            counters_o, _ = co.counters()
            counters_g, _ = cg.boundaryCounters()
            waiting, _, admitted = co.pending_state()
            print(f"production step {n + 1}: E, B, J rel L2 {errs}, admitted {admitted}, waiting "
                  f"{waiting}, window exits {counters_o[18]} / {counters_g[18]}", flush=True)
            assert max(errs.values()) <= 1e-12, errs
            assert np.array_equal(counters_o, counters_g)
    assert all(a == b for a, b in trials), trials
    waiting, _, admitted = co.pending_state()
    counters_o, _ = co.counters()
    # the bunch entered, released and began to leave through the exit plane
    assert admitted == queued and waiting == 0
    assert counters_o[18] > len(q) // 2
    xg, ug, _, _, _, idg = store.download()
    st = store.states()
    xo, uo = co.particles()
    ido = co.ids()
    live = st != 1
    og, oo = np.argsort(idg[live]), np.argsort(ido)
    assert np.array_equal(idg[live][og], ido[oo])
    released = st[live][og] == 0
    assert np.array_equal((~released).astype(np.int32), co.frozen()[oo])
    assert released.sum() > 0
    ref = xo[oo] - xo.mean(0)
    assert np.linalg.norm(xg[live][og] - xo.mean(0) - ref) / np.linalg.norm(ref) <= 1e-12
    assert np.linalg.norm(ug[live][og] - uo[oo]) / np.linalg.norm(uo) <= 1e-12
