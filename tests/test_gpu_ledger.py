"""Energy-ledger channels on the GPU against the oracle: the moving-window
field ledger (inlet overwrite, inserted, cropped and remap energies, the
trapezoid outward energy per face; moving_window_fields.cpp:207-218,
264-288, 387-438, 443-517) over whole coupled window steps, and the
particle sums of simulation.cpp:3591-3625. The bar is 1e-10 of the
channel's scale (FP64 compensated device sums against long-double host
sums of fields that agree to 1e-12); counts are exact."""
import numpy as np
import pytest

import harness as H
from test_gpu_coupling import case, planes_for

pytestmark = pytest.mark.gpu


def run_ledger_both(orc, co, gpu, win, cg, store, dims, steps, seed):
    rng = np.random.default_rng(seed)
    offset = 50
    origin = np.zeros(3)
    orc.enable_ledger()
    win.enableEnergyLedger()
    for n in range(steps):
        lower, inlet, shift = planes_for(rng, dims, range(3)), planes_for(rng, dims, range(2)), planes_for(rng, dims, range(6))
        t = n * 1e-15
        orc.clear_current(); gpu.clearCurrent()
        orc.window_begin(n, offset, lower); win.beginStep(n, offset, lower)
        orc.window_advance_magnetic(); gpu.advanceMagnetic()
        orc.window_set_inlet(offset, n + 1, *inlet); win.setInlet(*inlet)
        ro = co.step(origin, t)
        rg = cg.step(store, origin, t)
        assert ro[0] == rg[0]
        orc.finish_ledger_step(); win.finishEnergyLedgerStep()
        offset -= 1
        orc.window_shift(n + 1, offset, shift); win.shift(n + 1, offset, shift)
    return orc.ledger(), win.energyLedger()


def compare_ledger(lo, lg, tol):
    got = np.array([lg["inletOverwriteEnergyJ"], lg["insertedEnergyJ"], lg["croppedEnergyJ"], lg["remapEnergyJ"],
                    *lg["outwardEnergyFace"]])
    scale = np.abs(lo[:10]).max()
    err = np.abs(got - lo[:10])
    print("ledger", lo[:10], "max abs err / scale", err.max() / scale)
    assert np.all(err <= tol * np.maximum(np.abs(lo[:10]), 1e-3 * scale)), (got, lo[:10])
    assert [lg["overwriteEvents"], lg["fluxSteps"], lg["remapSteps"]] == list(lo[10:].astype(int))


@pytest.mark.parametrize("shape", ["binomial3", "quintic"])
def test_window_energy_ledger_matches_oracle(shape):
    dims = (20, 20, 40)
    orc, co, gpu, win, cg, store = case(dims, (4, 4, 0), 300, 17, shape=shape)
    lo, lg = run_ledger_both(orc, co, gpu, win, cg, store, dims, 4, 6)
    compare_ledger(lo, lg, 1e-10)


def test_particle_energy_sums_match_oracle():
    dims = (20, 20, 40)
    orc, co, gpu, win, cg, store = case(dims, (4, 4, 0), 300, 17)
    rng = np.random.default_rng(2)
    offset = 50
    for n in range(2):
        orc.clear_current(); gpu.clearCurrent()
        lower, inlet = planes_for(rng, dims, range(3)), planes_for(rng, dims, range(2))
        orc.window_begin(n, offset, lower); win.beginStep(n, offset, lower)
        orc.window_advance_magnetic(); gpu.advanceMagnetic()
        orc.window_set_inlet(offset, n + 1, *inlet); win.setInlet(*inlet)
        co.step(np.zeros(3), n * 1e-15)
        cg.step(store, np.zeros(3), n * 1e-15)
        so = co.particle_energy(n * 1e-15 + 3e-15)
        sg = cg.particleEnergy(store, n * 1e-15 + 3e-15)
        print(so, sg)
        assert np.all(np.abs(sg - so) <= 1e-12 * np.abs(so)), (sg - so) / so
        shift = planes_for(rng, dims, range(6))
        offset -= 1
        orc.window_shift(n + 1, offset, shift); win.shift(n + 1, offset, shift)


def test_ledger_phase_errors():
    import aprl_amd as A
    dims = (20, 20, 40)
    orc, co, gpu, win, cg, store = case(dims, (4, 4, 0), 16, 17)
    win.enableEnergyLedger()
    gpu.clearCurrent()
    win.beginStep(0, 50)
    gpu.advanceMagnetic()
    win.setInlet()
    cg.step(store, np.zeros(3), 0.0)
    with pytest.raises(A.AprlError) as e:
        win.shift(1, 49)
    assert "must finish after MPI and before shift" in str(e.value)
    win.finishEnergyLedgerStep()
    with pytest.raises(A.AprlError) as e:
        win.finishEnergyLedgerStep()
    assert "requires one completed MPI field step" in str(e.value)
    win.shift(1, 49)
    with pytest.raises(A.AprlError) as e:
        win.enableEnergyLedger()
    assert "before the first step" in str(e.value)
