"""The moving window's inlet reference on the GPU (aprl_inlet_reference_*,
aprl_window_set_reference) against the oracle's restatement of
BlockedCowanReference (oracle/orc_inlet_reference.hpp; pinned in
tests/test_oracle_inlet_reference.py against the Cowan initializer and the
27-point operator, since the reference unit needs FFTW3):

* the prepared E(0), B(-dt/2) for the CIC, linear-binomial3 + x-binomial3
  and quintic sources within 1e-12 of the oracle's (largest difference over
  the component's largest value; the charge is summed in fixed point and the
  transforms in another order, so not bitwise), the source charge exact, the
  mode eigenvalues bitwise (host arithmetic as the reference's);
* movingWindowFieldValue planes within 1e-12, the reference's error past the
  prepared planes;
* a window stepped with its planes taken from the device reference
  (lower halo at offset - 1, inlet at step + 1, inserted planes at the new
  offset) matches the oracle's window fed the oracle reference's planes to
  1e-12 of each vector field's largest value over 12 steps, and equals
  bitwise the same GPU window fed the device reference's own planes through
  the host.
"""
import numpy as np
import pytest

import harness as H
from test_oracle_inlet_reference import GUARDS, SPACING, OrcReference
from test_oracle_initial_field import bunch_case

pytestmark = pytest.mark.gpu

REF_DIMS = (24, 24, 96)
DT = SPACING[2] / H.C_SI
BETA = 0.29


def reference_pair(shape, x_binomial, seed=21):
    import aprl_amd as A
    x, u, q, m = bunch_case(REF_DIMS, SPACING, 400, 0.3, seed=seed)
    orc = OrcReference(REF_DIMS, SPACING, DT, np.zeros(3), x, q, BETA, GUARDS, shape=shape, x_binomial=x_binomial)
    g = A.EBGridGeometry(*REF_DIMS, *SPACING, DT)
    names = {"cic": "cic", "binomial3": "linear-binomial3", "quintic": "quintic-compensated"}
    gpu = A.BlockedCowanReference(g, np.zeros(3), x, q, BETA, GUARDS, names[shape], x_binomial)
    return orc, gpu, q


@pytest.mark.parametrize("shape,x_binomial", [("cic", False), ("binomial3", True), ("quintic", False)])
def test_reference_fields_match_oracle(shape, x_binomial):
    print(f"inlet reference {shape} x-binomial={x_binomial}", flush=True)
    orc, gpu, q = reference_pair(shape, x_binomial)
    for a in range(6):
        o, g = orc.component(a), gpu.field(a)
        scale = np.abs(o).max()
        err = np.abs(g - o).max() / scale if scale > 0 else np.abs(g).max()
        print(f"component {a}: {err:.2e}", flush=True)
        assert err <= 1e-12, (a, err)
    r = gpu.report
    assert r["sourceCharge"] == orc.report[2]
    assert r["depositedCharge"] == pytest.approx(orc.report[3], rel=1e-13)
    assert r["minimumModeEigenvalue"] == orc.report[0]
    assert r["minimumPivot"] == pytest.approx(orc.report[1], rel=1e-14)
    assert r["interiorNodes"] == tuple(n - 2 * g - 1 for n, g in zip(REF_DIMS, GUARDS))


def test_moving_planes_match_oracle():
    import aprl_amd as A
    orc, gpu, q = reference_pair("binomial3", True)
    for a in range(6):
        for k, step, tb in ((40, 0, 0.0), (40, 7, BETA), (1, 3, 0.3), (REF_DIMS[2] - 1, 0, 0.5)):
            o, g = orc.moving(a, k, step, tb), gpu.movingValue(a, k, step, tb)
            scale = max(np.abs(o).max(), 1e-300)
            assert np.abs(g - o).max() <= 1e-12 * scale, (a, k, step)
    with pytest.raises(A.OutOfRange, match="Moving reference query outside prepared native field"):
        gpu.movingValue(0, REF_DIMS[2] + 1, 0, 0.5)
    with pytest.raises(A.OutOfRange, match="outside prepared native field"):
        gpu.movingValue(2, 1, 5, 0.5)


def test_window_takes_its_planes_from_the_device_reference():
    import aprl_amd as A
    orc_ref, gpu_ref, q = reference_pair("binomial3", True)
    dims, cells = (REF_DIMS[0], REF_DIMS[1], 32), (3, 3, 0)
    offset0, steps = 40, 12
    # the window starts from the reference's own planes offset0 .. offset0 + nz (fill)
    orc = H.CpuFields("orc", *dims, *SPACING, DT, 1)
    orc.install_window(cells)
    win_fields = []
    for precision_case in ("reference", "host-planes"):
        gpu = A.EBFieldGrid(A.EBGridGeometry(*dims, *SPACING, DT))
        win = A.MovingWindowFields(gpu, A.EBCPMLParameters(cells))
        win_fields.append((gpu, win))
    for a in range(6):
        comp = orc_ref.component(a)
        nz_a = H.component_shape(*dims, a)[0]
        block = comp[offset0:offset0 + nz_a]
        orc.view(a)[...] = block
        for gpu, _ in win_fields:
            gpu.set_component(a, block)
    (gpu_a, win_a), (gpu_b, win_b) = win_fields
    win_a.setReference(gpu_ref, BETA)
    offset = offset0
    for n in range(steps):
        print(f"reference window step {n}", flush=True)
        lower = [orc_ref.moving(a, offset - 1, n, BETA) for a in range(3)]
        inlet = [orc_ref.moving(a, offset, n + 1, BETA) for a in range(2)]
        shift = [orc_ref.moving(a, offset - 1, n + 1, BETA) for a in range(6)]
        lower_g = [gpu_ref.movingValue(a, offset - 1, n, BETA) for a in range(3)]
        inlet_g = [gpu_ref.movingValue(a, offset, n + 1, BETA) for a in range(2)]
        shift_g = [gpu_ref.movingValue(a, offset - 1, n + 1, BETA) for a in range(6)]
        orc.clear_current(); gpu_a.clearCurrent(); gpu_b.clearCurrent()
        orc.window_begin(n, offset, lower); win_a.beginStep(n, offset); win_b.beginStep(n, offset, lower_g)
        orc.window_advance_magnetic(); gpu_a.advanceMagnetic(); gpu_b.advanceMagnetic()
        orc.window_set_inlet(offset, n + 1, *inlet); win_a.setInlet(); win_b.setInlet(*inlet_g)
        orc.advance_electric(); gpu_a.advanceElectric(); gpu_b.advanceElectric()
        offset -= 1
        orc.window_shift(n + 1, offset, shift); win_a.shift(n + 1, offset); win_b.shift(n + 1, offset, shift_g)
    # Bz starts at zero and grows only from the curl of the translating E,
    # where the larger components cancel: every component is measured
    # against the largest value of its vector field
    for comps in ((0, 1, 2), (3, 4, 5)):
        scale = max(np.abs(orc.view(a)).max() for a in comps)
        for a in comps:
            ga, gb, o = gpu_a.component(a), gpu_b.component(a), orc.view(a)
            assert np.array_equal(ga, gb), a
            assert np.abs(ga - o).max() <= 1e-12 * scale, (a, np.abs(ga - o).max() / scale)
