"""GPU charge-matched inlet vs the oracle.

Window steps with particles near the inlet, the centered normal reference
at every step start and the inlet completion after every shift: the plain
FP64 form matches the oracle to 1e-12 (the completion sums the inlet charge
in fixed point and its plane update in FP64 where the reference uses long
double); the FP32-split form tracks it with its FP32 error.
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


def run(shape, precision, steps=4, seed=6):
    import aprl_amd as A
    oshape, smooth, gshape = SHAPES[shape]
    orc, co = E.inlet_oracle_case("orc", oshape, smooth)
    gpu = A.EBFieldGrid(A.EBGridGeometry(*E.DIMS, E.DX, E.DX, E.DZ, E.DZ / H.C_SI), precision=precision)
    win = A.MovingWindowFields(gpu, A.EBCPMLParameters(E.CELLS))
    win.setChargeMatch(True)
    for a in range(9):
        gpu.set_component(a, orc.view(a))
    cg = A.ConservativeSourceCoupling(gpu, gshape)
    cg.setFrame(E.GAMMA_BOOST)
    cg.setParticleBoundary(E.CELLS)
    x, u, q, m, ids = E.inlet_particles()
    store = A.ParticleStore(gpu, len(q))
    store.upload(x, u, q, m, np.ones(len(q)), ids)
    rng = np.random.default_rng(seed)
    offset = E.OFFSET0
    ez_plane = E.planes_for(rng, [2])[0]
    for n in range(steps):
        print(f"inlet window step {n}", flush=True)
        origin = E.window_origin(offset)
        lower, inlet, shift = E.planes_for(rng, range(3)), E.planes_for(rng, range(2)), E.planes_for(rng, range(6))
        t = n * E.DZ / H.C_SI
        orc.clear_current(); gpu.clearCurrent()
        orc.window_begin_centered(n, offset, lower, ez_plane); win.beginStepCentered(n, offset, lower, ez_plane)
        orc.window_advance_magnetic(); gpu.advanceMagnetic()
        orc.window_set_inlet(offset, n + 1, *inlet); win.setInlet(*inlet)
        co.step(origin, t, 1)
        cg.step(store, origin, t)
        offset -= 1
        orc.window_shift(n + 1, offset, shift); win.shift(n + 1, offset, shift)
        co.complete_inlet(E.window_origin(offset))
        win.completeInlet(cg, store, E.window_origin(offset))
        co.finish_window(E.window_origin(offset))
        ez_plane = shift[2]
    return orc, gpu


@pytest.mark.parametrize("shape", ["binomial3", "quintic"])
def test_inlet_fp64_matches_oracle(shape):
    orc, gpu = run(shape, "fp64")
    for a in range(9):
        err = rel_l2(gpu.component(a), orc.view(a))
        assert err <= 1e-12, f"component {a}: rel L2 {err:.3e}"
    # the completed layer itself
    ez_g, ez_o = gpu.component(2), orc.view(2)
    assert rel_l2(ez_g[0], ez_o[0]) <= 1e-12


def test_inlet_fp32_split_tracks_oracle():
    orc, gpu = run("binomial3", "fp32-split")
    for a in range(6):
        err = rel_l2(gpu.component(a), orc.view(a))
        assert err <= 2e-5, f"component {a}: rel L2 {err:.3e}"
