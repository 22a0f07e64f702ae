"""Batched window steps of independent instances (aprl_window_step_batch,
aprl_window_run_batch) against the same instances stepped one by one with
the single-instance calls: fields, particles and trial counts must be
bitwise equal (the batch only interleaves the enqueue order of independent
streams), for both forms.
"""
import numpy as np
import pytest

from aprl_amd import workloads

pytestmark = pytest.mark.gpu


def small_instances(seed):
    ws = []
    for i, (dims, n) in enumerate([((32, 32, 256), 3000), ((32, 32, 256), 800), ((48, 48, 384), 5000)]):
        aperture = (dims[0] - 16) * 5.0e-6
        ws.append(workloads.Workload(
            name=f"test instance {i}", dims=dims, spacing=(5.0e-6, 5.0e-6, 2.0e-6), cpml=8, particles=n,
            gamma_lab=150.0 + 200 * i, gamma_boost=5.0, sigma=(aperture / 12, aperture / 12, 2e-6), sigma_gamma=5e-4,
            emittance=2e-7, charge_electrons=1e5, dipole_tesla=0.3 + 0.2 * i, dipole_length=0.2, seed=seed + i,
            z_cells=60.0))
    return ws


@pytest.mark.parametrize("precision", ["fp64", "fp32-split"])
def test_batch_equals_single_instance_steps(precision):
    ws = small_instances(91)
    batched = [(w, *workloads.setup_gpu(w, "linear-binomial3", precision)) for w in ws]
    single = [(w, *workloads.setup_gpu(w, "linear-binomial3", precision)) for w in ws]
    bs = workloads.BatchStepper(batched)
    steppers = [workloads.Stepper(w, f, win, cg, st) for (w, f, win, cg, st) in single]
    for _ in range(3):
        bs.step()
        for s in steppers:
            s.step()
    for b, s, (w, f1, _, _, st1), (_, f2, _, _, st2) in zip(range(len(ws)), steppers, batched, single):
        assert bs.trials[-1][b] == s.trials[-1]
        for a in range(9):
            assert np.array_equal(f1.component(a), f2.component(a)), (precision, b, a)
        x1, u1, *_r1, id1 = st1.download()
        x2, u2, *_r2, id2 = st2.download()
        assert np.array_equal(id1, id2) and np.array_equal(x1, x2) and np.array_equal(u1, u2), (precision, b)


@pytest.mark.parametrize("precision", ["fp64", "fp32-split"])
def test_run_batch_equals_single_instance_steps(precision):
    """aprl_window_run_batch: four steps in one call (the library advances
    step, offset, origin and time, and enqueues an instance's next step as
    soon as its own is committed) equal four single-instance steps bitwise,
    with the same summed trial counts."""
    ws = small_instances(37)
    batched = [(w, *workloads.setup_gpu(w, "linear-binomial3", precision)) for w in ws]
    single = [(w, *workloads.setup_gpu(w, "linear-binomial3", precision)) for w in ws]
    bs = workloads.BatchStepper(batched)
    steppers = [workloads.Stepper(w, f, win, cg, st) for (w, f, win, cg, st) in single]
    its = bs.run(4)
    for s in steppers:
        for _ in range(4):
            s.step()
    for b, s, (w, f1, _, _, st1), (_, f2, _, _, st2) in zip(range(len(ws)), steppers, batched, single):
        assert its[b] == sum(s.trials), (precision, b)
        assert bs.steps[b].step == 4 and bs.steps[b].offset == -4
        assert bs.steps[b].origin[2] == s.origin[2] and bs.steps[b].time == s.time
        for a in range(9):
            assert np.array_equal(f1.component(a), f2.component(a)), (precision, b, a)
        x1, u1, *_r1, id1 = st1.download()
        x2, u2, *_r2, id2 = st2.download()
        assert np.array_equal(id1, id2) and np.array_equal(x1, x2) and np.array_equal(u1, u2), (precision, b)


def test_batch_rejects_duplicate_instances():
    import aprl_amd as A
    w = small_instances(5)[1]
    inst = (w, *workloads.setup_gpu(w, "linear-binomial3", "fp32-split"))
    bs = workloads.BatchStepper([inst, inst])
    with pytest.raises(A.AprlError) as e:
        bs.step()
    assert "twice" in str(e.value)
