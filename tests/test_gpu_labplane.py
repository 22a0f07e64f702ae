"""Laboratory field planes on the GPU (aprl_lab_plane_*) against the
oracle's restatement of sampleFieldPlanes (lab_detectors.cpp:2426-2487):

* seeded fields, one grid in both: every sample bitwise equal in the plain
  FP64 form -- planes inside the grid, in the clamped end half-cells,
  beyond the grid (never sampled), and a lab-time rhythm of three steps
  (lab times and taken flags equal too);
* the FP32-split grid with the same fields within 1e-6 of the largest
  record per component (FP32 storage);
* the generated-Gaussian drift (32x32x256, 1e4 electrons, gamma_b = 10)
  stepped on both with a plane crossing the window: the GPU's records of
  every step within 1e-12 relative L2 of the oracle's, which samples the
  oracle's own fields (the fields agree to 1e-12 per step).
"""
import numpy as np
import pytest

import config_cases as CC
import harness as H
from aprl_amd import workloads as W
from test_oracle_labplane import OrcPlane, plane_time_for

pytestmark = pytest.mark.gpu


def grid_pair(dims, spacing, seed, precision="fp64"):
    import aprl_amd as A
    orc = H.CpuFields("orc", *dims, *spacing, spacing[2] / H.C_SI, 1)
    H.seeded_fields(orc, seed, components=range(6), scale=1e3)
    gpu = A.EBFieldGrid(A.EBGridGeometry(*dims, *spacing, spacing[2] / H.C_SI), precision=precision)
    for a in range(9):
        gpu.set_component(a, orc.view(a))
    cg = A.ConservativeSourceCoupling(gpu)
    return orc, gpu, cg


@pytest.mark.parametrize("precision", ["fp64", "fp32-split"])
def test_plane_samples_match_oracle(precision):
    import aprl_amd as A
    dims, spacing = (12, 10, 24), (2e-6, 2e-6, 1e-6)
    orc, gpu, cg = grid_pair(dims, spacing, 9, precision)
    gamma = 4.0
    cg.setFrame(gamma)
    frame = (gamma, 0.0, 0.0, 0.0)
    dt = spacing[2] / H.C_SI
    step_lab = dt / gamma
    origin = np.zeros(3)
    z_lab = 3e-5
    rhythm = 3 * step_lab * (1 - 1e-9)
    plane = A.LabFieldPlane(cg, z_lab, rhythm, capacity=16)
    ref = OrcPlane(z_lab, rhythm)
    # box times from the plane 0.1 cells above the grid down to below it, a
    # quarter step apart (the plane's box z falls by beta c dt per step)
    t0 = plane_time_for(frame, z_lab, (dims[2] + 0.1) * spacing[2])
    times_o, recs_o, flags = [], [], []
    for n in range(0, 110, 1):
        t = t0 + 0.25 * n * dt
        print(f"lab plane {precision} sample {n}", flush=True)
        taken_o, t_lab, rec = ref.sample(orc, frame, t, origin)
        taken_g = plane.sample(t, origin)
        assert taken_g == taken_o, (n, taken_g, taken_o)
        flags.append(taken_o)
        if taken_o:
            times_o.append(t_lab)
            recs_o.append(rec)
        if len(times_o) == 16 or n == 109:
            times_g, recs_g = plane.fetch()
            assert len(times_g) == len(times_o)
            assert np.array_equal(times_g, np.array(times_o))
            ro = np.array(recs_o)
            if precision == "fp64":
                assert np.array_equal(recs_g, ro)
            else:
                scale = np.abs(ro).max(axis=(0, 1, 2))
                assert np.all(np.abs(recs_g - ro) <= 1e-6 * scale), np.abs(recs_g - ro).max(axis=(0, 1, 2)) / scale
            times_o, recs_o = [], []
    # the plane entered the grid, was sampled on its rhythm and left it
    assert any(flags) and not flags[0] and not flags[-1]


def test_plane_through_coupled_drift_steps_matches_oracle():
    import aprl_amd as A
    w = W.DRIFT_C1
    data = w.generate()
    orc, co = CC.oracle_for(w, "linear-binomial3", data)
    fields, window, coupling, store = W.setup_gpu(w, "linear-binomial3", "fp64", data=data)
    gs = W.Stepper(w, fields, window, coupling, store)
    os_ = CC.OracleStepper(w, orc, co)
    frame = (w.gamma_boost, 0.0, 0.0, 0.0)
    z_lab = w.gamma_boost * 120 * w.spacing[2]  # box z = 120 cells at t = 0, crossing the window
    step_lab = w.dt / w.gamma_boost
    plane = A.LabFieldPlane(coupling, z_lab, 0.5 * step_lab, capacity=8)
    ref = OrcPlane(z_lab, 0.5 * step_lab)
    for n in range(4):
        print(f"coupled lab plane step {n}", flush=True)
        os_.step()
        gs.step()
        taken_o, t_lab, rec = ref.sample(orc, frame, os_.time, os_.origin)
        assert plane.sample(gs.time, gs.origin) == taken_o and taken_o
        t_g, rec_g = plane.fetch()
        assert t_g[0] == t_lab
        err = CC.rel_l2(rec_g[0], rec)
        print(f"step {n}: records rel L2 {err:.2e}", flush=True)
        assert err <= 1e-12
