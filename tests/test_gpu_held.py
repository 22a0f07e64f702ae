"""GPU held (frozen) particles vs the oracle.

Held electrons inside the window and upstream of it (beta_held = 0.99,
or ballistic along their momentum without the transverse hold) with an
activation plane that releases them during the run, next to active
electrons: the held motion, the upstream shaped support, the admission at
the inlet, the release kink with the push of the rest of the step, and the
residual of released particles. FP64 to 1e-12 with identical Picard trial
counts and held/released states; FP32-split with its FP32 error.
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


def run(shape, precision, steps=6, seed=12, hold=True):
    import aprl_amd as A
    oshape, smooth, gshape = SHAPES[shape]
    orc, co = E.held_oracle_case("orc", oshape, smooth, hold=hold)
    gpu = A.EBFieldGrid(A.EBGridGeometry(*E.DIMS, E.DX, E.DX, E.DZ, E.DZ / H.C_SI), precision=precision)
    win = A.MovingWindowFields(gpu, A.EBCPMLParameters(E.CELLS))
    for a in range(9):
        gpu.set_component(a, orc.view(a))
    cg = A.ConservativeSourceCoupling(gpu, gshape)
    cg.setFrame(E.GAMMA_BOOST)
    cg.setParticleBoundary(E.CELLS)
    cg.setActivationPlane(E.activation_lab_z())
    cg.setHoldTransverse(hold)
    x, u, q, m, ids, frozen, beta = E.held_particles()
    store = A.ParticleStore(gpu, len(q))
    store.upload(x, u, q, m, np.ones(len(q)), ids)
    store.setHeld(frozen, beta)
    rng = np.random.default_rng(seed)
    offset = E.OFFSET0
    logs = []
    for n in range(steps):
        print(f"held window step {n}", flush=True)
        origin = E.window_origin(offset)
        if n:
            co.finish_window(origin)
        lower, inlet, shift = E.planes_for(rng, range(3)), E.planes_for(rng, range(2)), E.planes_for(rng, range(6))
        t = n * E.DZ / H.C_SI
        orc.clear_current(); gpu.clearCurrent()
        orc.window_begin(n, offset, lower); win.beginStep(n, offset, lower)
        orc.window_advance_magnetic(); gpu.advanceMagnetic()
        orc.window_set_inlet(offset, n + 1, *inlet); win.setInlet(*inlet)
        ro = co.step(origin, t, 1)
        rg = cg.step(store, origin, t)
        logs.append((ro[0], rg[0]))
        offset -= 1
        orc.window_shift(n + 1, offset, shift); win.shift(n + 1, offset, shift)
    return orc, co, gpu, store, logs


def compare(orc, co, gpu, store, tol):
    for a in range(9):
        err = rel_l2(gpu.component(a), orc.view(a))
        assert err <= tol, f"component {a}: rel L2 {err:.3e}"
    xg, ug, _, _, _, idg = store.download()
    st = store.states()
    xo, uo = co.particles()
    ido = co.ids()
    fo = co.frozen()
    og, oo = np.argsort(idg), np.argsort(ido)
    assert np.array_equal(idg[og], ido[oo])
    assert np.array_equal((st[og] == 2).astype(np.int32), fo[oo])
    ref = xo[oo] - xo.mean(0)
    assert rel_l2(xg[og] - xo.mean(0), ref) <= tol
    assert rel_l2(ug[og], uo[oo]) <= tol
    return fo


@pytest.mark.parametrize("hold", [True, False])
@pytest.mark.parametrize("shape", ["binomial3", "quintic"])
def test_held_fp64_matches_oracle(shape, hold):
    orc, co, gpu, store, logs = run(shape, "fp64", hold=hold)
    assert all(a == b for a, b in logs), logs
    fo = compare(orc, co, gpu, store, 1e-12)
    x0, *_ , frozen0, beta = E.held_particles()
    assert fo.sum() > 0 and fo.sum() < frozen0.sum()  # some released, some still held


@pytest.mark.parametrize("hold", [True, False])
def test_held_fp32_split_tracks_oracle(hold):
    orc, co, gpu, store, logs = run("binomial3", "fp32-split", hold=hold)
    compare(orc, co, gpu, store, 2e-5)
