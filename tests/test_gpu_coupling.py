"""GPU conservative-source step vs the oracle (FP64 plain form).

The GPU sums currents tile by tile instead of particle by particle and keeps
the preview E* pristine instead of adding and removing each trial current,
so fields agree to round-off, not bitwise; the bar is 1e-12 relative L2 per
component and per particle state, with identical Picard trial counts.
"""
import numpy as np
import pytest

import harness as H
from aprl_amd import bunch

pytestmark = pytest.mark.gpu

DZ = 1.0e-6
TOL = 1e-12


def rel_l2(a, b):
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


SHAPES = {"binomial3": ("binomial3", True, "linear-binomial3"), "quintic": ("quintic", False, "quintic-compensated")}


def case(dims, cells, n_particles, seed, gamma_u=10.0, spread=(1.0, 1.0, 3.0), shape="binomial3", magnet="dipole"):
    import aprl_amd as A
    oshape, smooth, gshape = SHAPES[shape]
    dx = 2.5 * DZ
    orc = H.CpuFields("orc", *dims, dx, dx, DZ, DZ / H.C_SI, 1)
    orc.install_window(cells)
    H.seeded_fields(orc, seed, scale=1e5)
    co = H.CpuCoupling(orc, oshape, smooth, cpml_cells=cells)
    co.set_frame(5.0)
    # the lab element: a 0.5 T dipole, or a planar undulator / quadrupole
    # over the bunch's lab z (about 1e-4 m at t = 0)
    if magnet == "dipole":
        co.add_magnet("dipole", [0, 0, -1.0], 2.0, peak=0.5)
    elif magnet == "undulator":
        co.add_magnet("undulator", [0, 0, 0.0], 1.0, period=0.02, peak=1.2, angle=0.3, end_ramp=1.0, phase=0.4)
    else:
        co.add_magnet("quadrupole", [0, 0, 0.0], 1.0, gradient=40.0)
    g = A.EBGridGeometry(*dims, dx, dx, DZ, DZ / H.C_SI)
    gpu = A.EBFieldGrid(g)
    win = A.MovingWindowFields(gpu, A.EBCPMLParameters(cells))
    for a in range(9):
        gpu.set_component(a, orc.view(a))
    cg = A.ConservativeSourceCoupling(gpu, gshape)
    cg.setFrame(5.0)
    if magnet == "dipole":
        cg.addPrescribedLabMagnet("uniform-dipole", [0, 0, -1.0], 2.0, peak_field=0.5)
    elif magnet == "undulator":
        cg.addPrescribedLabMagnet("planar-undulator", [0, 0, 0.0], 1.0, period=0.02, peak_field=1.2,
                                  polarization_angle=0.3, end_ramp_periods=1.0, undulator_phase=0.4)
    else:
        cg.addPrescribedLabMagnet("quadrupole", [0, 0, 0.0], 1.0, quadrupole_gradient=40.0)
    cg.setParticleBoundary(cells)
    nx, ny, nz = dims
    center = [0.5 * nx * dx, 0.5 * ny * dx, 0.45 * nz * DZ]
    x, u, q, m, ids = bunch.seeded_gaussian(n_particles, seed, center,
                                            [spread[0] * dx, spread[1] * dx, spread[2] * DZ],
                                            [0.0, 0.0, gamma_u], [0.02, 0.02, 0.05], 3.0e5)
    co.set_particles(x, u, q, m, np.ones(n_particles), ids)
    store = A.ParticleStore(gpu, n_particles)
    store.upload(x, u, q, m, np.ones(n_particles), ids)
    return orc, co, gpu, win, cg, store


def planes_for(rng, dims, comps):
    return [rng.standard_normal(int(np.prod(H.component_shape(*dims, a)[1:]))) * (10.0 if a < 3 else 10.0 / H.C_SI)
            for a in comps]


def run_both(orc, co, gpu, win, cg, store, dims, steps, seed, check_each=False, substeps=1):
    rng = np.random.default_rng(seed)
    offset = 50
    origin = np.zeros(3)
    logs = []
    for n in range(steps):
        lower, inlet, shift = planes_for(rng, dims, range(3)), planes_for(rng, dims, range(2)), planes_for(rng, dims, range(6))
        t = n * 1e-15
        orc.clear_current(); gpu.clearCurrent()
        orc.window_begin(n, offset, lower); win.beginStep(n, offset, lower)
        orc.window_advance_magnetic(); gpu.advanceMagnetic()
        orc.window_set_inlet(offset, n + 1, *inlet); win.setInlet(*inlet)
        ro = co.step(origin, t, substeps)
        rg = cg.step(store, origin, t, substeps)
        logs.append((ro[0], rg[0]))
        if check_each:
            compare(orc, co, gpu, store)
        offset -= 1
        orc.window_shift(n + 1, offset, shift); win.shift(n + 1, offset, shift)
    return logs


def compare(orc, co, gpu, store, tol=TOL):
    for a in range(9):
        err = rel_l2(gpu.component(a), orc.view(a))
        assert err <= tol, f"component {a}: rel L2 {err:.3e}"
    for h in range(12):
        err = rel_l2(gpu.checkpointHistory(h), orc.history(h))
        assert err <= tol, f"history {h}: rel L2 {err:.3e}"
    xo, uo = co.particles()
    xg, ug, _, _, _, ids = store.download()
    order = np.argsort(ids)
    assert np.array_equal(np.sort(ids), np.arange(1, len(ids) + 1, dtype=np.uint64))
    assert rel_l2(xg[order] - xo.mean(0), xo - xo.mean(0)) <= tol
    assert rel_l2(ug[order], uo) <= tol


@pytest.mark.parametrize("shape", ["binomial3", "quintic"])
def test_coupled_steps_match_oracle(shape):
    dims = (20, 20, 40)
    args = case(dims, (4, 4, 0), 200, 17, shape=shape)
    logs = run_both(*args, dims, 4, 4, check_each=True)
    assert all(a == b for a, b in logs), logs
    assert all(a >= 2 for a, _ in logs)


@pytest.mark.parametrize("magnet,substeps", [("undulator", 4), ("quadrupole", 1)])
def test_prescribed_undulator_and_quadrupole_match_oracle(magnet, substeps):
    """SIMagneticElement::fieldsLab of the planar undulator (with its end
    ramps, phase and polarization; several push substeps per step, as the
    reference uses with an undulator) and of the quadrupole, boosted into the
    box frame, through whole coupled steps."""
    dims = (20, 20, 40)
    args = case(dims, (4, 4, 0), 200, 29, magnet=magnet)
    logs = run_both(*args, dims, 3, 6, check_each=True, substeps=substeps)
    assert all(a == b for a, b in logs), logs
    # the element's field acts on the bunch (box-frame B at a particle)
    xo, _ = args[1].particles()
    _, b = args[1].prescribed_box(xo[0], 0.0)
    assert np.linalg.norm(b) > 1e-3


@pytest.mark.parametrize("shape", ["binomial3", "quintic"])
def test_dense_bunch_many_particles_per_tile(shape):
    dims = (24, 24, 48)
    args = case(dims, (4, 4, 0), 4000, 23, spread=(0.8, 0.8, 2.0), shape=shape)
    logs = run_both(*args, dims, 3, 9)
    assert all(a == b for a, b in logs), logs
    compare(*args[:3], args[5])


def test_speculative_trial_batches_do_not_change_the_step():
    """The device decides convergence (simulation_source_coupling.cpp:46-51)
    and later trials of a batch return at once: one trial per host wait,
    eight enqueued ahead and the adaptive batch give the oracle's trial
    counts and bitwise the same fields and particles."""
    dims = (24, 24, 48)
    runs = []
    for batch in (1, 8, 0):
        args = case(dims, (4, 4, 0), 4000, 23, spread=(0.8, 0.8, 2.0))
        args[4].setTrialBatch(batch)
        print(f"trial batch {batch}", flush=True)
        logs = run_both(*args, dims, 3, 9)
        assert all(a == b for a, b in logs), (batch, logs)
        gpu, store = args[2], args[5]
        runs.append(([gpu.component(a) for a in range(9)], store.download()))
    for comps, parts in runs[1:]:
        for a in range(9):
            assert np.array_equal(comps[a], runs[0][0][a]), a
        for x, y in zip(parts, runs[0][1]):
            assert np.array_equal(x, y)


def test_gpu_step_is_deterministic_and_sort_is_stable():
    import aprl_amd as A
    dims = (20, 20, 40)
    results = []
    for rep in range(2):
        orc, co, gpu, win, cg, store = case(dims, (4, 4, 0), 300, 31)
        x0, _, _, _, _, ids0 = store.download()
        rng = np.random.default_rng(2)
        gpu.clearCurrent()
        win.beginStep(0, 50, planes_for(rng, dims, range(3)))
        gpu.advanceMagnetic()
        win.setInlet(*planes_for(rng, dims, range(2)))
        cg.step(store, np.zeros(3), 0.0)
        x, u, _, _, _, ids = store.download()
        # expected permutation: stable sort of the start-of-step tile keys
        dx = 2.5 * DZ
        cx = np.clip(np.floor(x0[:, 0] / dx).astype(int), 0, dims[0] - 1)
        cy = np.clip(np.floor(x0[:, 1] / dx).astype(int), 0, dims[1] - 1)
        cz = np.clip(np.floor(x0[:, 2] / DZ).astype(int), 0, dims[2] - 1)
        ntx, nty = -(-dims[0] // 8), -(-dims[1] // 8)
        key = ((cz // 4) * nty + cy // 8) * ntx + cx // 8
        expected = ids0[np.argsort(key, kind="stable")]
        assert np.array_equal(ids, expected)
        results.append((x, u, [gpu.component(a) for a in range(9)]))
    assert np.array_equal(results[0][0], results[1][0]) and np.array_equal(results[0][1], results[1][1])
    for a in range(9):
        assert np.array_equal(results[0][2][a], results[1][2][a])


def test_particle_starting_in_cpml_raises_out_of_range():
    """ParticleCPMLRegion::firstEntry's start check (particle_boundary.cpp:
    118-121) surfaces as the reference's out_of_range message."""
    import aprl_amd as A
    dims = (20, 20, 40)
    orc, co, gpu, win, cg, store = case(dims, (4, 4, 0), 16, 3)
    x, u, q, m, w, ids = store.download()
    x[0, 0] = 2.0 * 2.5 * DZ  # inside the x- CPML layer
    store.upload(x, u, q, m, w, ids)
    gpu.clearCurrent()
    win.beginStep(0, 50)
    gpu.advanceMagnetic()
    win.setInlet()
    with pytest.raises(A.AprlError) as e:
        cg.step(store, np.zeros(3), 0.0)
    assert "starts inside the non-physical CPML region" in str(e.value)
