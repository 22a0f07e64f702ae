"""The launch-shape autotuner (aprl_autotune, aprl_autotune_report): a run
that tunes the Cowan sweep's z-run length and the E update's workgroup size
on its first steps equals bitwise the same run with tuning off (every
candidate launch shape computes every node the same way), and the
decisions taken are reported per kernel and grid shape."""
import numpy as np
import pytest

import aprl_amd as A
from aprl_amd import workloads as W

pytestmark = pytest.mark.gpu


def instance(dims, seed):
    aperture = (dims[0] - 16) * 5.0e-6
    return W.Workload(name="autotune instance", dims=dims, spacing=(5.0e-6, 5.0e-6, 2.0e-6), cpml=8, particles=4000,
                      gamma_lab=200.0, gamma_boost=5.0, sigma=(aperture / 12, aperture / 12, 2e-6), sigma_gamma=5e-4,
                      emittance=2e-7, charge_electrons=1e5, dipole_tesla=0.3, dipole_length=0.2, seed=seed,
                      z_cells=60.0)


@pytest.mark.parametrize("precision", ["fp64", "fp32-split"])
def test_tuned_run_equals_untuned_run(precision):
    dims = (40, 40, 200) if precision == "fp64" else (40, 40, 208)  # shapes no other test tunes
    w = instance(dims, 17)
    runs = []
    try:
        for tuned in (False, True):
            A.setAutotune(tuned)
            f, win, cg, st = W.setup_gpu(w, "linear-binomial3", precision)
            s = W.Stepper(w, f, win, cg, st)
            for _ in range(16):
                s.step()
            f.synchronize()
            runs.append(([f.component(a) for a in range(9)], st.download(), s.trials))
    finally:
        A.setAutotune(True)
    (c0, p0, t0), (c1, p1, t1) = runs
    assert t0 == t1
    for a in range(9):
        assert np.array_equal(c0[a], c1[a]), a
    for x, y in zip(p0, p1):
        assert np.array_equal(x, y)
    word = "fp64" if precision == "fp64" else "fp32"
    shape = f"{dims[0]}x{dims[1]}x{dims[2]} {word}"
    decisions = A.autotuneDecisions()
    print(decisions, flush=True)
    assert decisions[f"cowan_march {shape}"] >= 8
    assert decisions[f"electric {shape}"] in (64, 128, 256)
