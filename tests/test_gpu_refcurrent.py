"""GPU reference surface current vs the oracle.

Coupled window steps with particles and the prescribed surface current at
the four transverse interfaces (a table linear in z plus a random part):
the ring's Jz is the trial's particle current plus the interpolated table
in every Picard trial and at the commit. FP64 form to 1e-12, FP32-split
form with its FP32 error.
"""
import numpy as np
import pytest

import event_cases as E
import harness as H

pytestmark = pytest.mark.gpu

SHAPES = {"binomial3": ("binomial3", True, "linear-binomial3"), "quintic": ("quintic", False, "quintic-compensated")}
BETA = 0.9375


def rel_l2(a, b):
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def run(shape, precision, steps=4, seed=8):
    import aprl_amd as A
    oshape, smooth, gshape = SHAPES[shape]
    orc, co = E.oracle_case("orc", oshape, smooth)
    ij = co.reference_points(E.CELLS[:2])
    rng = np.random.default_rng(seed)
    table = rng.normal(0, 1e4, (110, len(ij))) + np.arange(110)[:, None] * rng.normal(0, 1e3, len(ij))[None, :]
    co.set_reference_current(table, E.CELLS[:2], BETA)
    gpu = A.EBFieldGrid(A.EBGridGeometry(*E.DIMS, E.DX, E.DX, E.DZ, E.DZ / H.C_SI), precision=precision)
    win = A.MovingWindowFields(gpu, A.EBCPMLParameters(E.CELLS))
    for a in range(9):
        gpu.set_component(a, orc.view(a))
    cg = A.ConservativeSourceCoupling(gpu, gshape)
    cg.setFrame(E.GAMMA_BOOST)
    cg.setParticleBoundary(E.CELLS)
    cg.setReferenceCurrent(table, E.CELLS[:2], BETA)
    x, u, q, m, ids = E.particles()
    store = A.ParticleStore(gpu, len(q))
    store.upload(x, u, q, m, np.ones(len(q)), ids)
    offset = E.OFFSET0
    logs = []
    for n in range(steps):
        print(f"reference-current window step {n}", flush=True)
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
        if n == steps - 1:
            jz_o, jz_g = orc.view(8).copy(), gpu.component(8)
        offset -= 1
        orc.window_shift(n + 1, offset, shift); win.shift(n + 1, offset, shift)
    return orc, gpu, logs, jz_o, jz_g


@pytest.mark.parametrize("shape", ["binomial3", "quintic"])
def test_reference_current_fp64_matches_oracle(shape):
    orc, gpu, logs, jz_o, jz_g = run(shape, "fp64")
    assert all(a == b for a, b in logs), logs
    assert rel_l2(jz_g, jz_o) <= 1e-12
    for a in range(6):
        err = rel_l2(gpu.component(a), orc.view(a))
        assert err <= 1e-12, f"component {a}: rel L2 {err:.3e}"


def test_reference_current_fp32_split_tracks_oracle():
    orc, gpu, logs, jz_o, jz_g = run("binomial3", "fp32-split")
    assert rel_l2(jz_g, jz_o) <= 2e-5
    for a in range(6):
        err = rel_l2(gpu.component(a), orc.view(a))
        assert err <= 2e-5, f"component {a}: rel L2 {err:.3e}"
