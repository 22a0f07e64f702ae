"""Benchmark of the MI355X conservative-source PIC step.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--error-study]

One step = Simulation::solve's per-step order on the bending-radiation
workload (64x64x512 Cowan-z moving window, 8-cell CFS-CPML, 1e5 electrons,
0.5 T dipole, linear-binomial3 + x smoothing, FP64): clear J, magnetic
update, conservative-source step (preview, Picard trials with gather / push /
deposit / residual, commit), window shift. Prints one JSON line: the plain
FP64 form's throughput, with the quintic shape and the FP32-split form of
both shapes timed the same way beside it. --error-study prints the
FP32-split form's errors against FP64 at 100, 1000 and 10000 steps instead.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TUNE_STEPS = 12
WORD = 8  # FP64

# Measured HBM bytes per launch of the plain FP64 form on this workload:
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, one counter per run, calibrated
# on a plain 256 MiB device copy (bytes = 2.0 x FETCH_SIZE + 1.0 x
# WRITE_SIZE, in KiB). Source and commands:
# profiles/bending_64x64x512_fp64_binomial3x_kernels_traffic.md. Traffic
# above the algorithmic bytes is re-reading (staged halos, per-chunk boxes).
MEASURED_TRAFFIC_FP64 = {
    "magnetic": (2 * 63341 + 66930 + 2 * 10527 + 1792) * 1024,
    "electric": (2 * 81838 + 64850) * 1024,
    "electric_preview": (2 * 22206 + 23402) * 1024,
    "gather": (2 * 10400 + 4689) * 1024,
    "residual": (2 * 13926 + 4733) * 1024,
    "deposit": (2 * 4317 + 26442) * 1024,
    "push": (2 * 5840 + 5859) * 1024,
    "reduce_current": (2 * 24171 + 46501) * 1024,
}


def kernel_bytes(w, name, tiles, chunks, active=0):
    """Algorithmic HBM bytes of one launch: what the kernel must read or
    write at least once. Fields: 9 words/cell magnetic (read 3 E, read and
    write 3 B), 12 words/cell electric (read 3 B, 3 J, read and write 3 E),
    plus 2 words per active CPML history entry; the zero-current preview
    12 words per node of each active 8x8x4 tile (read 3 B, 3 E; write E*
    and E* + E_old). Particles: push 14 words/particle (read x, u, q, m, 6
    fields; write x, u); gathers 14 words/particle plus 9 words per cell of
    each occupied 8x8x4 tile (E*, E_old, B); deposit 7 words/particle (x0,
    x1, q) plus 3 words per node of each occupied tile's 13x11x9 extended
    box."""
    cells = w.cells
    n = w.particles
    cp = w.cpml
    nx, ny, nz = w.dims
    psi = 2 * cp * (ny + nx) * nz * 2  # x and y slabs: two faces, two histories per field component pair
    tile_cells = 8 * 8 * 4
    ext_nodes = 13 * 11 * 9
    if name == "magnetic":
        return WORD * (9 * cells + 2 * psi)
    if name == "electric":
        return WORD * (12 * cells + 2 * psi)
    if name == "electric_preview":
        return WORD * 12 * tile_cells * active
    if name == "push":
        return WORD * 14 * n
    if name in ("gather", "residual"):
        return WORD * (14 * n + 9 * tile_cells * tiles)
    if name == "deposit":
        return WORD * (7 * n + 3 * ext_nodes * tiles)
    return None


def run_gpu(args, shape="linear-binomial3", profile=True, precision="fp64"):
    from aprl_amd import workloads
    w = workloads.BENDING
    fields, window, coupling, store = workloads.setup_gpu(w, shape, precision)
    s = workloads.Stepper(w, fields, window, coupling, store)
    s.prepare = not getattr(args, "sort_inline", False)
    # untimed steps in which the launch-shape autotuner times its candidates
    # for this grid (decided before the warm-up whatever its length)
    for _ in range(TUNE_STEPS):
        s.step()
    for _ in range(args.warmup):
        s.step()
    fields.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.step()
    fields.synchronize()
    elapsed = time.perf_counter() - t0
    trials = s.trials[args.warmup:]
    counts = (coupling.last_tiles, coupling.last_chunks, coupling.last_active_tiles)
    if not profile:
        return w, elapsed, trials, {}, counts
    # Per-kernel HIP-event durations (on the library's own stream) over a
    # further profiled window of the same steps, with the sort enqueued in
    # line (in the timed steps it overlaps the field updates on a second
    # stream, which would fold the two kernels' sharing of the device into
    # each one's duration).
    s.prepare = False
    fields.profile(True)
    prof_steps = max(5, min(args.steps, 20))
    for _ in range(prof_steps):
        s.step()
    fields.synchronize()
    kernels = {}
    for name in ("magnetic", "electric_preview", "electric", "shift", "sort", "gather", "push", "deposit",
                 "reduce_current", "residual", "clear_current"):
        launches, ms = fields.profile_query(name)
        if launches:
            kernels[name] = {"launches": launches, "avg_us": 1e3 * ms / launches, "total_ms": ms}
    fields.profile(False)
    return w, elapsed, trials, kernels, counts


def run_batch(count=48, steps=20, warmup=2):
    """The batched scan: `count` independent instances (workloads.
    batch_workloads) advanced together by one aprl_window_run_batch call per
    run of steps, in both forms from the same generated bunches;
    whole-batch particle-steps/s."""
    from aprl_amd import workloads
    ws = workloads.batch_workloads(count)
    data = [w.generate() for w in ws]
    out = {"instances": count, "particles": int(sum(w.particles for w in ws)), "steps": steps}
    for precision in ("fp32-split", "fp64"):
        inst = [(w, *workloads.setup_gpu(w, "linear-binomial3", precision, d)) for w, d in zip(ws, data)]
        bs = workloads.BatchStepper(inst)
        bs.run(warmup)
        for _, f, *_r in inst:
            f.synchronize()
        t0 = time.perf_counter()
        bs.run(steps)
        for _, f, *_r in inst:
            f.synchronize()
        elapsed = time.perf_counter() - t0
        out[precision.replace("-", "_")] = {"value": out["particles"] * steps / elapsed,
                                            "ms_per_step": 1e3 * elapsed / steps}
        del bs, inst
        print(f"batch {precision}: {1e3 * elapsed / steps:.2f} ms per step", file=sys.stderr, flush=True)
    out["fp32_split_vs_fp64"] = round(out["fp32_split"]["value"] / out["fp64"]["value"], 3)
    return out


def run_batch_production(count=48, steps=1000, capacity=100):
    """The batched scan as production runs: the same `count` instances
    (workloads.batch_workloads), each from plane-arrival events through the
    input route and its pending queue (workloads.batch_injection: the
    bending card's planes on its own grid), advanced together for `steps`
    steps by aprl_window_run_batch (queue admission, window step, lab field
    plane per step), in both forms. particle-steps count the store's
    entries at every step."""
    import numpy as np
    from aprl_amd import workloads
    ws = workloads.batch_workloads(count)
    out = {"workload": f"batched scan of {count} production runs (grids 32x32x256 to 64x64x512, "
                       "2e4-1e6 e-, dipole 0.1-1 T, gamma 100-1000)",
           "instances": count, "particles": int(sum(w.particles for w in ws)), "steps": steps}
    for precision in ("fp32-split", "fp64"):
        runs = [workloads.ProductionRun(w, workloads.batch_injection(w), "linear-binomial3", precision,
                                        max_steps=steps, plane_capacity=capacity) for w in ws]
        bp = workloads.BatchProduction(runs)
        for run in runs:
            run.fields.synchronize()
        t0 = time.perf_counter()
        bp.run(steps)
        for run in runs:
            run.fields.synchronize()
        elapsed = time.perf_counter() - t0
        ps = sum(run.particle_steps for run in runs)
        exits = sum(int(run.coupling.boundaryCounters()[0][18]) for run in runs)
        admitted = sum(int(run.queue.admittedCount()) for run in runs)
        trials = float(np.mean(bp.iterations)) / steps
        out[precision.replace("-", "_")] = {"ms_per_step": 1e3 * elapsed / steps, "seconds": round(elapsed, 3),
                                            "particle_steps": int(ps), "value": ps / elapsed,
                                            "unit": "particle-steps/s", "admitted": admitted,
                                            "window_exits": exits, "picard_trials_mean": round(trials, 3),
                                            "lab_plane_samples": int(sum(run.plane_samples for run in runs))}
        print(f"batched production {precision}: {1e3 * elapsed / steps:.2f} ms per step", file=sys.stderr,
              flush=True)
        del bp, runs
    out["fp32_split_vs_fp64"] = round(out["fp32_split"]["value"] / out["fp64"]["value"], 3)
    return out


def split_error(steps=100, shape="linear-binomial3"):
    """Relative L2 error of the FP32-split form against the plain FP64 form
    (itself within 1e-12 of the oracle per step) after `steps` steps of the
    same workload from the same state: E and B (as vector fields) and u."""
    import numpy as np
    from aprl_amd import workloads
    w = workloads.BENDING
    runs = []
    for precision in ("fp64", "fp32-split"):
        fields, window, coupling, store = workloads.setup_gpu(w, shape, precision)
        s = workloads.Stepper(w, fields, window, coupling, store)
        for _ in range(steps):
            s.step()
        x, u, _, _, _, ids = store.download()
        order = np.argsort(ids)
        runs.append(([fields.component(a) for a in range(6)], u[order], x[order]))
        del s, fields, window, coupling, store

    (f64, u64, x64), (f32, u32, x32) = runs
    return {"steps": steps, "e": field_error(f32[:3], f64[:3]), "b": field_error(f32[3:], f64[3:]),
            "u": field_error([u32], [u64]), "x_cells": float(np.abs((x32 - x64) / np.array(w.spacing)).max())}


def field_error(test, ref):
    """Relative L2 error of a vector field given as component arrays:
    |test - ref| / |ref| over all components together (a component that is
    zero up to round-off in the reference does not dominate)."""
    import numpy as np
    num = sum(float(np.sum((t - r) ** 2)) for t, r in zip(test, ref))
    den = sum(float(np.sum(r ** 2)) for r in ref)
    return float(np.sqrt(num / den)) if den > 0 else float(np.sqrt(num))


def error_study(horizons=(100, 1000, 10000), shape="linear-binomial3"):
    """FP32-split against the plain FP64 form on the co-moving drift bunch
    (stays inside the window for 10^4 steps) at each horizon: relative L2
    error of E and B (vector fields), of u, the largest position
    difference in cells, and of the power spectrum of Ex on the grid plane
    through the bunch (sampled every step, |rfft over time|^2 summed over
    the plane)."""
    import numpy as np
    from aprl_amd import workloads
    w = workloads.COMOVING
    forms = ("fp64", "fp32-split")
    setups = {p: workloads.setup_gpu(w, shape, p) for p in forms}
    steppers = {p: workloads.Stepper(w, *setups[p]) for p in forms}
    k0 = w.dims[2] // 2
    planes = {p: [] for p in forms}

    rows = []
    n = 0
    t0 = time.perf_counter()
    for horizon in horizons:
        while n < horizon:
            for p in forms:
                steppers[p].step()
                planes[p].append(setups[p][0].component_plane(0, k0))
            n += 1
        f64, f32 = setups["fp64"][0], setups["fp32-split"][0]
        us = []
        for p in forms:
            x, u, _, _, _, ids = setups[p][3].download()
            order = np.argsort(ids)
            us.append((x[order], u[order]))
        spectra = [np.sum(np.abs(np.fft.rfft(np.array(planes[p]), axis=0)) ** 2, axis=(1, 2)) for p in forms]
        c32 = [f32.component(a) for a in range(6)]
        c64 = [f64.component(a) for a in range(6)]
        rows.append({"steps": n, "e": field_error(c32[:3], c64[:3]), "b": field_error(c32[3:], c64[3:]),
                     "u": field_error([us[1][1]], [us[0][1]]),
                     "x_cells": float(np.abs((us[1][0] - us[0][0]) / np.array(w.spacing)).max()),
                     "plane_spectrum": field_error([spectra[1]], [spectra[0]]),
                     "components": [field_error([c32[a]], [c64[a]]) for a in range(6)]})
        print(f"error study: {n} steps ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    return {"workload": w.name, "shape": shape, "plane_k": k0, "rows": rows}


def host_cores():
    """Host CPU facts for the baseline record: physical cores and sockets
    (lscpu -p=CORE,SOCKET), the CPUs this process may run on, the model
    name, and the thread count used: every physical core this process may
    use, capped by OMP_NUM_THREADS where the environment sets it."""
    import subprocess
    info = {"affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=20).stdout
        pairs = {tuple(line.split(",")[:2]) for line in out.splitlines() if line and not line.startswith("#")}
        info["physical_cores"] = len(pairs)
        info["sockets"] = len({p[1] for p in pairs})
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    threads = min(info.get("physical_cores", info["affinity_cpus"]), info["affinity_cpus"])
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit() and int(cap) > 0:
        info["omp_num_threads_env"] = int(cap)
        threads = min(threads, int(cap))
    info["threads"] = max(1, threads)
    return info


def run_cpu_baseline(warmup=20, max_timed=40, seconds_budget=20.0):
    """The reference's own compiled hot path (oracle/_ref: its eb_field,
    eb_cpml, moving_window_fields, source_coupling and longitudinal_shape
    translation units with OpenMP) on the bending-radiation workload, in a
    child process whose OpenMP runtime starts with OMP_NUM_THREADS = the
    thread count of host_cores() and OMP_PROC_BIND=close, OMP_PLACES=cores.
    Timed over steps `warmup` -> N (the reference's own convention,
    verification/performance/cases/memory_scaling/README.md:43): N stops
    at `max_timed` timed steps or when the timed steps pass the budget."""
    import subprocess
    info = host_cores()
    env = dict(os.environ)
    env.update({"OMP_NUM_THREADS": str(info["threads"]), "OMP_PROC_BIND": "close", "OMP_PLACES": "cores"})
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-child", str(warmup), str(max_timed),
           str(seconds_budget)]
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600)
    if r.returncode != 0:
        print(f"cpu baseline failed: {r.stderr[-2000:]}", file=sys.stderr, flush=True)
        return None
    out = json.loads(r.stdout.strip().splitlines()[-1])
    if out is None:
        return None
    out["host"] = info
    return out


def cpu_baseline_child(warmup, max_timed, seconds_budget):
    import ctypes as C
    import statistics
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import harness as H
    from aprl_amd import workloads
    if not H.have_ref():
        print(json.dumps(None))
        return
    w = workloads.BENDING
    threads = int(os.environ["OMP_NUM_THREADS"])
    lib = H.load_particles("ref")
    lib.ref_set_threads.argtypes = [C.c_int]
    if lib.ref_set_threads(threads):
        print(json.dumps(None))
        return
    nx, ny, nz = w.dims
    dx, dy, dz = w.spacing
    f = H.CpuFields("ref", nx, ny, nz, dx, dy, dz, w.dt, 1)
    f.install_window((w.cpml, w.cpml, 0))
    c = H.CpuCoupling(f, "binomial3", True, cpml_cells=(w.cpml, w.cpml, 0))
    c.set_frame(w.gamma_boost)
    c.add_magnet("dipole", [0.0, 0.0, -0.5 * w.dipole_length], w.dipole_length, peak=w.dipole_tesla,
                 angle=0.5 * np.pi)
    x, u, q, m, ids = w.generate()
    c.set_particles(x, u, q, m, np.ones(len(q)), ids)
    zero3 = [np.zeros(int(np.prod(H.component_shape(nx, ny, nz, a)[1:]))) for a in range(3)]
    zero6 = [np.zeros(int(np.prod(H.component_shape(nx, ny, nz, a)[1:]))) for a in range(6)]
    offset = 10_000
    origin = np.zeros(3)
    times = []
    step = 0
    while True:
        t0 = time.perf_counter()
        f.clear_current()
        f.window_begin(step, offset, zero3)
        f.window_advance_magnetic()
        f.window_set_inlet(offset, step + 1, zero3[0], zero3[1])
        c.step(origin, step * w.dt)
        offset -= 1
        f.window_shift(step + 1, offset, zero6)
        origin[2] = (offset - 10_000) * dz
        dt_step = time.perf_counter() - t0
        if step >= warmup:
            times.append(dt_step)
        step += 1
        if len(times) >= 3 and (sum(times) > seconds_budget or len(times) >= max_timed):
            break
    per_step = sum(times) / len(times)
    third = len(times) // 3
    segments = [sum(times[i * third:(i + 1) * third]) / third for i in range(3)]
    print(json.dumps({
        "value": w.particles / per_step, "unit": "particle-steps/s", "cores": threads, "kind": "reference",
        "cell_updates_per_s": w.cells / per_step, "ms_per_step": 1e3 * per_step,
        "ms_per_step_median_of_3": 1e3 * statistics.median(segments),
        "sample": f"steps {warmup} -> {warmup + len(times)} of the bending-radiation workload (64x64x512, 1e5 e-, "
                  f"linear-binomial3 + x-binomial3, FP64), the first {warmup} untimed; reference TUs (eb_field, "
                  f"eb_cpml, moving_window_fields, source_coupling, longitudinal_shape) with OpenMP on {threads} "
                  f"threads, OMP_PROC_BIND=close, OMP_PLACES=cores"}))


def run_large(w, steps=3, warmup=1, data=None):
    """The FEL segment (128x128x1024, 4e6 quiet-start macro-particles) or the
    capacity case (256x256x2048, 6.4e7 macro-particles drawn on the device)
    in the FP32-split form with the quintic shape: `warmup` untimed steps,
    then `steps` timed ones (a shortened run of the configurations' 20000
    and 2000 steps)."""
    from aprl_amd import workloads
    fields, window, coupling, store = workloads.setup_gpu(w, "quintic-compensated", "fp32-split", data=data)
    coupling.setSplitTrials(2)
    s = workloads.Stepper(w, fields, window, coupling, store)
    for _ in range(warmup):
        s.step()
    fields.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step()
    fields.synchronize()
    elapsed = time.perf_counter() - t0
    out = {"workload": w.name, "particles": int(w.particles), "grid": list(w.dims), "form": "FP32-split",
           "shape": "quintic-compensated", "steps": steps, "warmup": warmup,
           "value": w.particles * steps / elapsed, "unit": "particle-steps/s",
           "cell_updates_per_s": w.cells * steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
           "occupied_tiles": coupling.last_tiles, "chunks": coupling.last_chunks}
    print(f"{w.name}: {1e3 * elapsed / steps:.1f} ms per step", file=sys.stderr, flush=True)
    del s, fields, window, coupling, store
    return out


def run_production(steps=2000, precision="fp64", shape="linear-binomial3", keep_patch=(16, 48, 16, 48)):
    """The bending-radiation card run as specified: 2000 steps from
    plane-arrival events through the input route and the pending queue,
    with the dipole, the lab exit plane, the energy ledger and a lab field
    plane sampled every step (workloads.ProductionRun). Timed over the whole
    run (queue admissions and plane samples included); particle-steps count
    every store entry (held ones too) at each step."""
    import numpy as np
    from aprl_amd import workloads
    w = workloads.BENDING
    run = workloads.ProductionRun(w, workloads.BENDING_INJECTION, shape, precision, max_steps=steps,
                                  keep_patch=keep_patch)
    run.fields.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run.step()
    run.fields.synchronize()
    elapsed = time.perf_counter() - t0
    counts, _ = run.coupling.boundaryCounters()
    ledger = run.window.energyLedger()
    times, records = run.finish()
    out = {"workload": w.name + ", frozen lab-plane injection", "form": "plain FP64" if precision == "fp64"
           else "FP32-split", "shape": shape, "steps": steps, "seconds": round(elapsed, 3),
           "ms_per_step": 1e3 * elapsed / steps, "particle_steps": int(run.particle_steps),
           "value": run.particle_steps / elapsed, "unit": "particle-steps/s",
           "cell_updates_per_s": w.cells * steps / elapsed, "queued_initially": int(run.queued0),
           "admitted": int(run.queue.admittedCount()), "window_exits": int(counts[18]),
           "cropped_carriers": int(counts[19]), "picard_trials_mean": sum(run.trials) / len(run.trials),
           "lab_plane_samples": int(run.plane_samples),
           "ledger_outward_energy_J": float(np.sum(ledger["outwardEnergyFace"])),
           "ledger_cropped_energy_J": float(ledger["croppedEnergyJ"])}
    print(f"production run ({out['form']}): {out['ms_per_step']:.3f} ms per step, {out['window_exits']} exits",
          file=sys.stderr, flush=True)
    del run
    return out, (times, records)


def plane_error(test, ref):
    """FP32-split against FP64 on the lab field plane of the production run
    (the central 32 x 32 nodes, every step): relative L2 of the E_lab and
    B_lab records as vector fields over the whole run, and of the power
    spectrum over lab time of E_x and B_y summed over the patch."""
    import numpy as np
    (tt, rt), (tr, rr) = test, ref
    n = min(len(tt), len(tr))
    rt, rr = rt[:n], rr[:n]
    out = {"samples": int(n)}
    for name, comps in (("e", (0, 1, 2)), ("b", (3, 4, 5))):
        out[name] = float(np.linalg.norm(rt[..., comps] - rr[..., comps]) / np.linalg.norm(rr[..., comps]))
    for name, c in (("ex_spectrum", 0), ("by_spectrum", 4)):
        st = np.abs(np.fft.rfft(rt[..., c], axis=0)) ** 2
        sr = np.abs(np.fft.rfft(rr[..., c], axis=0)) ** 2
        out[name] = float(np.linalg.norm(st.sum(axis=(1, 2)) - sr.sum(axis=(1, 2))) /
                          np.linalg.norm(sr.sum(axis=(1, 2))))
    return out


def production_snapshots(precision, at=(100, 1000, 2000), shape="linear-binomial3"):
    """E and B of the grid and (id, u) of the store after each step count in
    `at`, from an untimed production run of the bending card."""
    import numpy as np
    from aprl_amd import workloads
    run = workloads.ProductionRun(workloads.BENDING, workloads.BENDING_INJECTION, shape, precision,
                                  max_steps=max(at))
    shots = {}
    for n in range(1, max(at) + 1):
        run.step()
        if n in at:
            x, u, q, m, w, ids = run.store.download()
            order = np.argsort(ids)
            shots[n] = ([run.fields.component(a) for a in range(6)], ids[order], u[order])
    del run
    return shots


def production_accuracy():
    """FP32-split against FP64 on the bending card after 100, 1000 and 2000
    steps: relative L2 of E and B (as vector fields) on the grid, and of u
    over the entries present in both stores (matched by id)."""
    import numpy as np
    s64 = production_snapshots("fp64")
    s32 = production_snapshots("fp32-split")
    out = {}
    f0 = s64[min(s64)][0]
    first_norm = float(np.sqrt(sum(float(np.sum(f0[c] ** 2)) for c in range(3))))
    for n in sorted(s64):
        (f64, id64, u64), (f32, id32, u32) = s64[n], s32[n]
        err = {}
        for name, comps in (("e", (0, 1, 2)), ("b", (3, 4, 5))):
            num = sum(float(np.sum((f32[c] - f64[c]) ** 2)) for c in comps)
            den = sum(float(np.sum(f64[c] ** 2)) for c in comps)
            err[name] = float(np.sqrt(num / den)) if den > 0 else None
        # the field's size at this step beside its size at the first
        # snapshot (after the bunch has left, what stays in the window is
        # small, and its relative error larger)
        err["field_norm_vs_first"] = float(np.sqrt(sum(float(np.sum(f64[c] ** 2)) for c in range(3))) /
                                           first_norm) if first_norm else None
        common, i64, i32 = np.intersect1d(id64, id32, return_indices=True)
        if len(common):
            du = u32[i32] - u64[i64]
            err["u"] = float(np.linalg.norm(du) / np.linalg.norm(u64[i64]))
        err["entries"] = int(len(common))
        out[str(n)] = err
    return out


def production_legs():
    p64, plane64 = run_production()
    p32, plane32 = run_production(precision="fp32-split")
    p32["plane_error_vs_fp64"] = plane_error(plane32, plane64)
    p32["error_vs_fp64_by_step"] = production_accuracy()
    return {"fp64": p64, "fp32_split": p32}


def step_roofline(w, ms, word=WORD):
    """Whole-step bytes against HBM peak: 21 words per cell-update (B: read
    3 E, read and write 3 B; E: read 3 B, 3 J, read and write 3 E) plus 2
    words per CPML history entry of each update, and 14 words per
    particle-step (read x, u, q, m; write x, u)."""
    nx, ny, nz = w.dims
    psi = 2 * w.cpml * (ny + nx) * nz * 2
    bytes_cells = word * (21 * w.cells + 2 * 2 * psi)
    bytes_particles = word * 14 * w.particles
    achieved = (bytes_cells + bytes_particles) / (ms * 1e-3) / 1e9
    return {"bytes_per_step": bytes_cells + bytes_particles, "bytes_per_cell_update": 21 * word,
            "bytes_per_particle_step": 14 * word, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4)}


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--cpu-baseline-child":
        cpu_baseline_child(int(sys.argv[2]), int(sys.argv[3]), float(sys.argv[4]))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-large", action="store_true", help="skip the FEL segment and capacity legs")
    ap.add_argument("--leg", choices=("all", "fp64", "fp32-split", "production", "batch", "batch-production"),
                    default="all",
                    help="time only this form's binomial3-x leg (for profiling runs), only the 2000-step "
                         "production runs, only the batched scan, or only its 1000-step production form")
    ap.add_argument("--sort-inline", action="store_true",
                    help="with --leg: enqueue each step's sort in line instead of beside the field updates "
                         "(for per-kernel profiles without two kernels sharing the device)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="launch every kernel's default shape (for profiles whose averages should not include "
                         "the tuner's trial launches)")
    ap.add_argument("--error-study", action="store_true",
                    help="print the FP32-split vs FP64 errors at 100, 1000 and 10000 steps and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="three plain 256 MiB device-to-device copies (the byte reference of the HBM "
                         "traffic counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 32 * 1024 * 1024  # doubles: 256 MiB
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 8, "copies": 3}))
        return
    if args.no_autotune:
        import aprl_amd
        aprl_amd.setAutotune(False)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the whole window lives on one device); --gpus {args.gpus} "
              "is not supported", file=sys.stderr)
        sys.exit(2)
    if args.error_study:
        print(json.dumps({"error_study": error_study()}))
        return
    if args.leg == "production":
        print(json.dumps({"production": production_legs()}))
        return
    if args.leg == "batch":
        print(json.dumps({"batch": run_batch()}))
        return
    if args.leg == "batch-production":
        print(json.dumps({"batch_production": run_batch_production()}))
        return
    if args.leg != "all":
        w, elapsed, trials, _, (tiles, chunks, active) = run_gpu(args, profile=False, precision=args.leg)
        print(json.dumps({"leg": args.leg, "value": w.particles * args.steps / elapsed, "unit": "particle-steps/s",
                          "ms_per_step": 1e3 * elapsed / args.steps, "steps": args.steps, "warmup": args.warmup,
                          "occupied_tiles": tiles, "chunks": chunks, "active_tiles": active}))
        return
    w, elapsed, trials, kernels, (tiles, chunks, active) = run_gpu(args)
    # the other production shape of the same workload, timed the same way
    _, q_elapsed, q_trials, _, _ = run_gpu(args, "quintic-compensated", profile=False)
    # the FP32-split form of both shapes
    _, s_elapsed, s_trials, s_kernels, _ = run_gpu(args, precision="fp32-split")
    _, sq_elapsed, _, _, _ = run_gpu(args, "quintic-compensated", profile=False, precision="fp32-split")
    s_error = split_error()
    batch = run_batch()
    batch_production = run_batch_production()
    production = production_legs()
    large = None
    if not args.no_large:
        from aprl_amd import workloads
        large = {"fel_segment": run_large(workloads.FEL_C3),
                 "capacity": run_large(workloads.CAPACITY_C5, steps=2, data="device")}
    ms = 1e3 * elapsed / args.steps
    value = w.particles * args.steps / elapsed
    nbytes = {k: kernel_bytes(w, k, tiles, chunks, active) for k in kernels}
    modeled = {k: v for k, v in kernels.items() if nbytes[k] is not None}
    dominant = max(modeled.items(), key=lambda kv: kv[1]["total_ms"])[0] if modeled else None
    roofline = None
    if dominant:
        k = kernels[dominant]
        b = nbytes[dominant]
        achieved = b / (k["avg_us"] * 1e-6) / 1e9
        roofline = {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": MEASURED_TRAFFIC_FP64.get(dominant), "algorithmic_bytes": b,
                    "avg_launch_us": round(k["avg_us"], 2)}
    cpu = None if args.no_cpu_baseline else run_cpu_baseline()
    line = {
        "metric": "particle-steps/s + cell-updates/s, FP32-split vs FP64, % of HBM peak",
        "value": value,
        "unit": "particle-steps/s",
        "cell_updates_per_s": w.cells * args.steps / elapsed,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": w.name, "grid": list(w.dims), "particles": w.particles, "form": "plain FP64",
                   "shape": "linear-binomial3 + x-binomial3", "picard_trials_mean": sum(trials) / len(trials),
                   "occupied_tiles": tiles, "active_tiles": active, "chunks": chunks},
        "step_roofline": step_roofline(w, ms),
        "quintic": {"value": w.particles * args.steps / q_elapsed, "ms_per_step": 1e3 * q_elapsed / args.steps,
                    "picard_trials_mean": sum(q_trials) / len(q_trials)},
        "fp32_split": {"value": w.particles * args.steps / s_elapsed, "ms_per_step": 1e3 * s_elapsed / args.steps,
                       "vs_fp64": round(elapsed / s_elapsed, 3), "picard_trials": s_trials[0],
                       "quintic_value": w.particles * args.steps / sq_elapsed,
                       "quintic_vs_fp64": round(q_elapsed / sq_elapsed, 3),
                       "error_vs_fp64": s_error,
                       "step_roofline": step_roofline(w, 1e3 * s_elapsed / args.steps, word=4),
                       "kernels_us": {k: round(v["avg_us"], 2) for k, v in s_kernels.items()}},
        "batch": batch,
        "batch_production": batch_production,
        "autotune": __import__("aprl_amd").autotuneDecisions(),
        "production": production,
        "large": large,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels_us": {k: round(v["avg_us"], 2) for k, v in kernels.items()},
        "kernels_gbs": {k: round(nbytes[k] / (v["avg_us"] * 1e-6) / 1e9, 1) for k, v in modeled.items()},
        "kernels_traffic_gbs": {k: round(MEASURED_TRAFFIC_FP64[k] / (v["avg_us"] * 1e-6) / 1e9, 1)
                                for k, v in kernels.items() if k in MEASURED_TRAFFIC_FP64},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
