"""GPU particle-boundary events vs the oracle.

Window steps whose particles enter the transverse CPML, leave through the
upper face or pass the moving-window exit plane, so that carriers are made,
move, deposit damped current, leave and get cropped. The plain FP64 form
matches the oracle to 1e-12 (fields, particles, carriers, counter charges)
with identical counts and Picard trial counts; the FP32-split form matches
the same events with its FP32 field and state error.

The device crops the carriers of a shifted window when the next step starts
(against that step's origin, the one finishMovingWindowStep uses), so the
oracle applies its window finish at the same point.
"""
import numpy as np
import pytest

import event_cases as E
import harness as H

pytestmark = pytest.mark.gpu

SHAPES = {"binomial3": ("binomial3", True, "linear-binomial3"), "quintic": ("quintic", False, "quintic-compensated")}


def rel_l2(a, b):
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def gpu_case(shape, exit_plane, precision):
    import aprl_amd as A
    oshape, smooth, gshape = SHAPES[shape]
    orc, co = E.oracle_case("orc", oshape, smooth, exit_plane)
    gpu = A.EBFieldGrid(A.EBGridGeometry(*E.DIMS, E.DX, E.DX, E.DZ, E.DZ / H.C_SI), precision=precision)
    win = A.MovingWindowFields(gpu, A.EBCPMLParameters(E.CELLS))
    for a in range(9):
        gpu.set_component(a, orc.view(a))
    cg = A.ConservativeSourceCoupling(gpu, gshape)
    cg.setFrame(E.GAMMA_BOOST)
    cg.setParticleBoundary(E.CELLS, E.exit_lab_z() if exit_plane else float("inf"))
    x, u, q, m, ids = E.particles(exit_plane=exit_plane)
    store = A.ParticleStore(gpu, len(q))
    store.upload(x, u, q, m, np.ones(len(q)), ids)
    return orc, co, gpu, win, cg, store


def run_both(orc, co, gpu, win, cg, store, steps, exit_plane, seed=4):
    rng = np.random.default_rng(seed)
    offset = E.OFFSET0
    exit_z = E.exit_lab_z() if exit_plane else float("inf")
    logs = []
    for n in range(steps):
        print(f"event window step {n}", flush=True)
        origin = E.window_origin(offset)
        if n:
            co.finish_window(origin)
        lower, inlet, shift = E.planes_for(rng, range(3)), E.planes_for(rng, range(2)), E.planes_for(rng, range(6))
        t = n * E.DZ / H.C_SI
        orc.clear_current(); gpu.clearCurrent()
        orc.window_begin(n, offset, lower); win.beginStep(n, offset, lower)
        orc.window_advance_magnetic(); gpu.advanceMagnetic()
        orc.window_set_inlet(offset, n + 1, *inlet); win.setInlet(*inlet)
        ro = co.step(origin, t, 1, exit_z)
        rg = cg.step(store, origin, t)
        logs.append((ro[0], rg[0]))
        offset -= 1
        orc.window_shift(n + 1, offset, shift); win.shift(n + 1, offset, shift)
    return logs


def compare(orc, co, gpu, cg, store, tol_field, tol_state, exact_counts=True):
    for a in range(9):
        err = rel_l2(gpu.component(a), orc.view(a))
        assert err <= tol_field, f"component {a}: rel L2 {err:.3e}"
    xg, ug, qg, mg, wg, idg = store.download()
    st = store.states()
    assert len(st) == len(idg)
    # active particles
    xo, uo = co.particles()
    ido = co.ids()
    act = st == 0
    assert np.array_equal(np.sort(idg[act]), np.sort(ido))
    og = np.argsort(idg[act])
    oo = np.argsort(ido)
    if len(ido):
        ref = xo[oo] - xo.mean(0)
        assert rel_l2(xg[act][og] - xo.mean(0), ref) <= tol_state
        assert rel_l2(ug[act][og], uo[oo]) <= tol_state
    # carriers
    xc, vc, qc, wc, idc = co.carriers()
    car = st == 1
    assert np.array_equal(np.sort(idg[car]), np.sort(idc))
    cgo = np.argsort(idg[car])
    coo = np.argsort(idc)
    if len(idc):
        assert rel_l2(xg[car][cgo] - xc.mean(0), xc[coo] - xc.mean(0)) <= tol_state
        assert np.max(np.abs(wg[car][cgo] - wc[coo])) <= tol_state
        gamma = np.sqrt(1.0 + np.sum(ug[car][cgo] ** 2, axis=1))
        assert rel_l2(ug[car][cgo] * (H.C_SI / gamma)[:, None], vc[coo]) <= tol_state
    # counters
    cn_o, cq_o = co.counters()
    cn_g, cq_g = cg.boundaryCounters()
    if exact_counts:
        assert np.array_equal(cn_o, cn_g), (cn_o, cn_g)
    assert rel_l2(cq_g, cq_o) <= tol_state
    return cn_o


@pytest.mark.parametrize("shape", ["binomial3", "quintic"])
@pytest.mark.parametrize("exit_plane", [False, True])
def test_events_fp64_match_oracle(shape, exit_plane):
    case = gpu_case(shape, exit_plane, "fp64")
    logs = run_both(*case, 8, exit_plane)
    assert all(a == b for a, b in logs), logs
    orc, co, gpu, win, cg, store = case
    counts = compare(orc, co, gpu, cg, store, 1e-12, 1e-12)
    assert counts[0] > 0 and counts[3] > 0
    if exit_plane:
        assert counts[18] > 0 and counts[17] > 0 and counts[19] > 0
    else:
        assert counts[11] > 0


@pytest.mark.parametrize("shape", ["binomial3", "quintic"])
def test_events_fp32_split_match_oracle(shape):
    case = gpu_case(shape, True, "fp32-split")
    case[4].setSplitTrials(3)
    run_both(*case, 8, True)
    orc, co, gpu, win, cg, store = case
    counts = compare(orc, co, gpu, cg, store, 2e-5, 2e-5)
    assert counts[0] > 0 and counts[18] > 0


def test_removed_entries_leave_the_store():
    """Counts after every step: live entries = oracle particles + carriers,
    and a second run of the same steps gives the same device state."""
    runs = []
    for _ in range(2):
        case = gpu_case("binomial3", True, "fp64")
        orc, co, gpu, win, cg, store = case
        n0 = len(store)
        run_both(*case, 5, True)
        assert len(store) == len(co.ids()) + len(co.carriers()[4])
        assert len(store) < n0
        runs.append((store.download(), store.states(), cg.boundaryCounters(), [gpu.component(a) for a in range(9)]))
    (d0, s0, c0, f0), (d1, s1, c1, f1) = runs
    for a, b in zip(d0, d1):
        assert np.array_equal(a, b)
    assert np.array_equal(s0, s1)
    assert np.array_equal(c0[0], c1[0]) and np.array_equal(c0[1], c1[1])
    for a, b in zip(f0, f1):
        assert np.array_equal(a, b)
