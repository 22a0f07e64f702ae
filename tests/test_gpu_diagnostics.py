"""GPU FP64 diagnostics: energy-ledger channels, charge continuity and the
Gauss-law residual.

Field energy and face power are compared with long-double sums of the
oracle's per-cell fieldEnergyDensityCell / poyntingComponentCell (pinned
bitwise to the reference build) to 1e-13 relative. Continuity: after a
coupled step the device's (rho1 - rho0)/dt + div J must vanish to 1e-12 of
the charge rate. Gauss: with fields built Gauss-consistent from the shaped
charge, div E - rho/eps0 stays at round-off (1e-12 of the source) after a
coupled step in the interior.
"""
import numpy as np
import pytest

import harness as H
from aprl_amd import bunch
from test_gpu_coupling import SHAPES, planes_for

pytestmark = pytest.mark.gpu

DZ = 1.0e-6


def test_energy_and_face_power_match_long_double_sums():
    import aprl_amd as A
    dims = (14, 12, 18)
    dx, dy = 2 * DZ, 1.5 * DZ
    orc = H.CpuFields("orc", *dims, dx, dy, DZ, DZ / H.C_SI, 1)
    H.seeded_fields(orc, 8, scale=1e4)
    gpu = A.EBFieldGrid(A.EBGridGeometry(*dims, dx, dy, DZ, DZ / H.C_SI))
    for a in range(9):
        gpu.set_component(a, orc.view(a))
    lo, hi = (3, 2, 4), (11, 10, 15)
    vol = np.longdouble(dx) * np.longdouble(dy) * np.longdouble(DZ)
    ref = np.longdouble(0)
    for k in range(lo[2], hi[2]):
        for j in range(lo[1], hi[1]):
            for i in range(lo[0], hi[0]):
                ref += np.longdouble(orc.energy_density(i, j, k)) * vol
    got = gpu.interiorFieldEnergy(lo, hi)
    assert abs(got - float(ref)) <= 1e-13 * abs(float(ref))
    area = [np.longdouble(dy) * DZ, np.longdouble(dx) * DZ, np.longdouble(dx) * dy]
    power = np.zeros(6, dtype=np.longdouble)
    for axis in range(3):
        u, v = [a for a in range(3) if a != axis]
        for side in range(2):
            c = [0, 0, 0]
            c[axis] = hi[axis] - 1 if side else lo[axis]
            for b in range(lo[v], hi[v]):
                for a in range(lo[u], hi[u]):
                    c[u], c[v] = a, b
                    s = np.longdouble(orc.poynting(c[0], c[1], c[2], axis)) * area[axis]
                    power[2 * axis + side] += s if side else -s
    got = gpu.facePower(lo, hi)
    scale = float(np.abs(power).max())
    assert np.all(np.abs(got - power.astype(float)) <= 1e-13 * scale), (got, power)


def coupled_case(shape, seed, dims=(20, 20, 40), cells=(4, 4, 0), n=300, precision="fp64"):
    import aprl_amd as A
    _, _, gshape = SHAPES[shape]
    dx = 2.5 * DZ
    gpu = A.EBFieldGrid(A.EBGridGeometry(*dims, dx, dx, DZ, DZ / H.C_SI), precision=precision)
    win = A.MovingWindowFields(gpu, A.EBCPMLParameters(cells))
    cg = A.ConservativeSourceCoupling(gpu, gshape)
    cg.setFrame(5.0)
    cg.addPrescribedLabMagnet("uniform-dipole", [0, 0, -1.0], 2.0, peak_field=0.5)
    cg.setParticleBoundary(cells)
    nx, ny, nz = dims
    x, u, q, m, ids = bunch.seeded_gaussian(n, seed, [0.5 * nx * dx, 0.5 * ny * dx, 0.45 * nz * DZ],
                                            [dx, dx, 3 * DZ], [0.0, 0.0, 10.0], [0.02, 0.02, 0.05], 3.0e5)
    store = A.ParticleStore(gpu, n)
    store.upload(x, u, q, m, np.ones(n), ids)
    return gpu, win, cg, store, (x, q), dx


def one_step(gpu, win, cg, store, rng, dims):
    gpu.clearCurrent()
    win.beginStep(0, 50, planes_for(rng, dims, range(3)))
    gpu.advanceMagnetic()
    win.setInlet(*planes_for(rng, dims, range(2)))
    cg.step(store, np.zeros(3), 0.0)


@pytest.mark.parametrize("shape", ["binomial3", "quintic"])
def test_coupled_step_continuity_on_device(shape):
    dims = (20, 20, 40)
    gpu, win, cg, store, _, _ = coupled_case(shape, 41)
    rng = np.random.default_rng(4)
    one_step(gpu, win, cg, store, rng, dims)
    residual, rate, _ = cg.continuity(store)
    assert rate > 0
    assert residual <= 1e-12 * rate, (residual, rate)


@pytest.mark.parametrize("shape", ["binomial3", "quintic"])
def test_gauss_residual_stays_at_round_off(shape):
    dims = (20, 20, 40)
    gpu, win, cg, store, (x, q), dx = coupled_case(shape, 43)
    h = (dx, dx, DZ)
    # Gauss-consistent start: Ez accumulates rho along z, Ex = Ey = 0, B = 0
    rho = H.shaped_charge(shape, x, q, dims, h)
    ez = np.zeros((dims[2], dims[1] + 1, dims[0] + 1))
    acc = np.zeros((dims[1] + 1, dims[0] + 1))
    for k in range(dims[2]):
        acc = acc + rho[k] * DZ / H.EPS0
        ez[k] = acc
    gpu.set_component(2, ez)
    lo, hi = (6, 6, 4), (14, 14, 36)
    before = cg.gaussResidual(store, np.zeros(3), lo, hi)
    assert before[2] <= 1e-12 * before[3], before
    rng = np.random.default_rng(5)
    one_step(gpu, win, cg, store, rng, dims)
    after = cg.gaussResidual(store, np.zeros(3), lo, hi)
    assert after[3] > 0
    assert after[2] <= 1e-12 * after[3], after
