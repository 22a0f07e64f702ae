"""Synthetic workloads of the benchmark configurations.

``bending_radiation`` is the single-GPU headline case: 1e5 unit-weight
electrons from a seeded 6D Gaussian (lab gamma 100, sigma_x = sigma_y =
20 um, sigma_z = 5 um, sigma_gamma/gamma = 5e-4, normalized emittance 1 um)
in the gamma_b = 5 boosted frame, a 64 x 64 x 512 Cowan-z moving window with
8-cell transverse CFS-CPML, a 0.5 T lab dipole over 0.2 m, and the
linear-binomial3 (+x) source shape. The bunch is an active box-frame
snapshot (its lab sample is boosted with the light-front transform and its
length stretched to the box frame), the inlet reference is the zero field,
and no particle reaches a boundary during the timed steps.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import bunch

C_SI = 299792458.0


@dataclass
class Workload:
    name: str
    dims: tuple
    spacing: tuple
    cpml: int
    particles: int
    gamma_lab: float
    gamma_boost: float
    sigma: tuple            # lab sigma_x, sigma_y, sigma_z [m]
    sigma_gamma: float      # relative
    emittance: float        # normalized [m]
    charge_electrons: float  # represented electrons in total
    dipole_tesla: float
    dipole_length: float
    seed: int
    z_cells: float          # bunch centre, cells above the window origin
    box_uz: float = 0.0     # nonzero: box-frame mean u_z of a bunch given directly in the box frame

    @property
    def dt(self) -> float:
        return self.spacing[2] / C_SI

    @property
    def cells(self) -> int:
        return int(np.prod(self.dims))

    def generate(self):
        """Box-frame particles: position (n,3), u (n,3), charge, mass, ids."""
        nx, ny, nz = self.dims
        dx, dy, dz = self.spacing
        if self.box_uz:
            center = [0.5 * nx * dx, 0.5 * ny * dy, self.z_cells * dz]
            sig_u = [self.emittance / self.sigma[0], self.emittance / self.sigma[1], self.sigma_gamma * abs(self.box_uz)]
            return bunch.seeded_gaussian(self.particles, self.seed, center, list(self.sigma), [0.0, 0.0, self.box_uz],
                                         sig_u, self.charge_electrons)
        u_lab = np.sqrt(self.gamma_lab ** 2 - 1.0)
        gb = bunch.gamma_from_u(np.array([[0.0, 0.0, u_lab]]))  # lab gamma
        ub = bunch.lab_to_box_u(np.array([[0.0, 0.0, u_lab]]), self.gamma_boost)
        g_box = bunch.gamma_from_u(ub)[0]
        stretch = gb[0] / g_box  # length of a moving bunch in the box frame
        center = [0.5 * nx * dx, 0.5 * ny * dy, self.z_cells * dz]
        sig_pos = [self.sigma[0], self.sigma[1], self.sigma[2] * stretch]
        sig_u = [self.emittance / self.sigma[0], self.emittance / self.sigma[1], self.sigma_gamma * self.gamma_lab]
        x, u, q, m, ids = bunch.seeded_gaussian(self.particles, self.seed, center, sig_pos, [0.0, 0.0, u_lab], sig_u,
                                                self.charge_electrons)
        u = bunch.lab_to_box_u(u, self.gamma_boost)
        return x, u, q, m, ids


BENDING = Workload(
    name="bending-radiation 64x64x512 dipole, 1e5 e-",
    dims=(64, 64, 512), spacing=(5.0e-6, 5.0e-6, 2.0e-6), cpml=8, particles=100_000,
    gamma_lab=100.0, gamma_boost=5.0, sigma=(20e-6, 20e-6, 5e-6), sigma_gamma=5e-4, emittance=1e-6,
    charge_electrons=100_000.0, dipole_tesla=0.5, dipole_length=0.2, seed=20241, z_cells=150.0)


# Precision study: a drift-only bunch that moves with the window. The window
# moves by -dz per step (c dt = dz); a bunch with box-frame u_z = -30 trails
# it by 1/(2 u_z^2) = 5.6e-4 cells per step, so it stays inside for 10^4
# steps (no particle boundary events) while its fields evolve.
COMOVING = Workload(
    name="co-moving drift 32x32x256, 1e4 e-",
    dims=(32, 32, 256), spacing=(5.0e-6, 5.0e-6, 2.0e-6), cpml=8, particles=10_000,
    gamma_lab=30.0, gamma_boost=10.0, sigma=(7.5e-6, 7.5e-6, 10e-6), sigma_gamma=1e-3, emittance=7.5e-9,
    charge_electrons=6.0e5, dipole_tesla=0.0, dipole_length=0.0, seed=777, z_cells=128.0, box_uz=-30.0)


def setup_gpu(w: Workload, shape: str = "linear-binomial3", precision: str = "fp64", data=None):
    """Field grid + window + coupling + particle store for workload w;
    shape is "linear-binomial3" (with x-binomial3 smoothing) or
    "quintic-compensated", precision "fp64" or "fp32-split"; data = the
    output of w.generate() when already generated."""
    import aprl_amd as A
    g = A.EBGridGeometry(*w.dims, *w.spacing, w.dt)
    fields = A.EBFieldGrid(g, precision=precision)
    window = A.MovingWindowFields(fields, A.EBCPMLParameters((w.cpml, w.cpml, 0)))
    coupling = A.ConservativeSourceCoupling(fields, shape)
    coupling.setFrame(w.gamma_boost)
    # lab dipole B_y over z_lab in [-L/2, L/2]
    if w.dipole_tesla:
        coupling.addPrescribedLabMagnet("uniform-dipole", [0.0, 0.0, -0.5 * w.dipole_length], w.dipole_length,
                                        peak_field=w.dipole_tesla, polarization_angle=0.5 * np.pi)
    coupling.setParticleBoundary((w.cpml, w.cpml, 0))
    x, u, q, m, ids = w.generate() if data is None else data
    store = A.ParticleStore(fields, len(q))
    store.upload(x, u, q, m, None, ids)
    return fields, window, coupling, store


class Stepper:
    """Drives Simulation::solve's per-step order (simulation.cpp:518-583):
    clear J, window begin, magnetic, coupled source step, shift."""

    def __init__(self, w: Workload, fields, window, coupling, store, offset: int = 10_000):
        self.w, self.fields, self.window, self.coupling, self.store = w, fields, window, coupling, store
        self.step_index = 0
        self.offset = offset
        self.origin = np.zeros(3)
        self.time = 0.0
        self.trials = []

    def step(self):
        f, win = self.fields, self.window
        n = self.step_index
        f.clearCurrent()
        win.beginStep(n, self.offset)
        f.advanceMagnetic()
        win.setInlet()
        its, _, _ = self.coupling.step(self.store, self.origin, self.time)
        self.trials.append(its)
        self.offset -= 1
        win.shift(n + 1, self.offset)
        # Simulation::finishMovingWindowStep: origin = fma(offset, dz, ref)
        # with the reference origin at offset 10000 (fma(a, b, 0) = a*b).
        self.origin[2] = (self.offset - 10_000) * self.w.spacing[2]
        self.time += self.w.dt
        self.step_index += 1


class BatchStepper:
    """Steps several independent instances (each a workload with its own
    grid, window, coupling and store) with aprl_window_step_batch: the
    per-instance bookkeeping of Stepper, one C-ABI call per step."""

    def __init__(self, instances, offset: int = 10_000):
        import ctypes as C
        from . import _capi
        self.instances = instances  # (workload, fields, window, coupling, store)
        self.steps = (_capi.InstanceStep * len(instances))()
        for s, (w, f, win, cg, st) in zip(self.steps, instances):
            s.fields, s.coupling, s.particles = f.handle, cg.handle, st.handle
            s.step, s.offset, s.time, s.substeps = 0, offset, 0.0, 1
            for a in range(3):
                s.origin[a] = 0.0
        self._C = C
        self.trials = []

    def step(self):
        from ._capi import check, lib
        check(lib().aprl_window_step_batch(self.steps, len(self.instances)))
        self.trials.append([s.report.iterations for s in self.steps])
        for s, (w, *_rest) in zip(self.steps, self.instances):
            s.step += 1
            s.offset -= 1
            # Simulation::finishMovingWindowStep: origin = fma(offset, dz, ref)
            s.origin[2] = (s.offset - 10_000) * w.spacing[2]
            s.time += w.dt


def batch_workloads(count: int, seed: int = 4242):
    """Independent instances of the batched scan: particle counts
    log-uniform in [2e4, 1e6], grids drawn from 32x32x256, 48x48x384 and
    64x64x512 (8-cell transverse CPML), lab gamma uniform in [100, 1000], a
    lab dipole uniform in [0.1, 1] T; transverse sizes scaled to each grid's
    physical aperture so that no particle starts in the CPML."""
    rng = np.random.default_rng(seed)
    grids = [(32, 32, 256), (48, 48, 384), (64, 64, 512)]
    out = []
    for i in range(count):
        dims = grids[int(rng.integers(0, 3))]
        n = int(round(np.exp(rng.uniform(np.log(2e4), np.log(1e6)))))
        gamma = float(rng.uniform(100.0, 1000.0))
        tesla = float(rng.uniform(0.1, 1.0))
        aperture = (dims[0] - 2 * 8) * 5.0e-6
        sig = aperture / 12.0
        out.append(Workload(
            name=f"batch instance {i}", dims=dims, spacing=(5.0e-6, 5.0e-6, 2.0e-6), cpml=8, particles=n,
            gamma_lab=gamma, gamma_boost=5.0, sigma=(sig, sig, 2e-6), sigma_gamma=5e-4, emittance=1e-6 * sig / 20e-6,
            charge_electrons=float(n), dipole_tesla=tesla, dipole_length=0.2, seed=seed + 17 * i,
            z_cells=60.0))
    return out
