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

The other configurations: ``DRIFT_C1`` (the CPU-runnable generated-Gaussian
drift, 32 x 32 x 256), ``fel_segment`` (the planar-undulator FEL segment,
128 x 128 x 1024, 4e6 quiet-start macro-particles of weight 355), the
batched scan ``batch_workloads`` (48 instances) and ``CAPACITY_C5`` (6.4e7
macro-particles in 256 x 256 x 2048 with a 16-cell CPML).
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
    undulator: tuple = ()   # (K, period [m], segments, periods per segment, break [m]) of a planar undulator line
    substeps: int = 1       # particle substeps per field step

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

    def beam_parameters(self):
        """(center, sigma_position, mean_u, sigma_u, boost gamma) of the
        lab-frame draw that ``generate`` boosts into the box frame."""
        nx, ny, nz = self.dims
        dx, dy, dz = self.spacing
        center = [0.5 * nx * dx, 0.5 * ny * dy, self.z_cells * dz]
        if self.box_uz:
            sig_u = [self.emittance / self.sigma[0], self.emittance / self.sigma[1],
                     self.sigma_gamma * abs(self.box_uz)]
            return center, list(self.sigma), [0.0, 0.0, self.box_uz], sig_u, 1.0
        u_lab = np.sqrt(self.gamma_lab ** 2 - 1.0)
        gb = bunch.gamma_from_u(np.array([[0.0, 0.0, u_lab]]))
        g_box = bunch.gamma_from_u(bunch.lab_to_box_u(np.array([[0.0, 0.0, u_lab]]), self.gamma_boost))[0]
        sig_pos = [self.sigma[0], self.sigma[1], self.sigma[2] * gb[0] / g_box]
        sig_u = [self.emittance / self.sigma[0], self.emittance / self.sigma[1], self.sigma_gamma * self.gamma_lab]
        return center, sig_pos, [0.0, 0.0, u_lab], sig_u, self.gamma_boost

    def generate_on_device(self, store):
        """The same beam as ``generate`` drawn on the device straight into
        ``store`` (aprl_particles_generate_gaussian; the device's log and cos
        instead of numpy's)."""
        center, sig_pos, mean_u, sig_u, boost = self.beam_parameters()
        store.generateGaussian(self.particles, center, sig_pos, mean_u, sig_u, self.charge_electrons, self.seed,
                               boost_gamma=boost)


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


# Plain FP64, CPU-runnable: the generated-Gaussian drift (one electron bunch
# of 1e4 macro-particles, lab gamma 200, sigma_x = sigma_y = 30 um,
# sigma_z = 10 um, normalized emittance 1 um, sigma_gamma/gamma = 1e-3,
# 50 pC) in the gamma_b = 10 frame, 32 x 32 x 256 with 8 transverse CPML
# cells, no magnets. The 20 um transverse cells put the CPML 5.3 sigma from
# the axis; the box-frame bunch (sigma_z 200 um = 25 cells) is centred in
# the window.
DRIFT_C1 = Workload(
    name="generated-Gaussian drift 32x32x256, 1e4 e-",
    dims=(32, 32, 256), spacing=(20.0e-6, 20.0e-6, 8.0e-6), cpml=8, particles=10_000,
    gamma_lab=200.0, gamma_boost=10.0, sigma=(30e-6, 30e-6, 10e-6), sigma_gamma=1e-3, emittance=1e-6,
    charge_electrons=50e-12 / bunch.QE, dipole_tesla=0.0, dipole_length=0.0, seed=17, z_cells=128.0)


# Single-GPU capacity: 6.4e7 macro-particles of a seeded 6D Gaussian (lab
# gamma 500, sigma_x = sigma_y = 40 um, sigma_z = 30 um) in 256 x 256 x 2048
# with 16 transverse CPML cells, the planar undulator line of the FEL
# segment, quintic shape, FP32-split. gamma_b = 4 stretches the bunch to a
# box-frame sigma_z of 0.24 mm (118 cells): the farthest of 6.4e7 Gaussian
# draws (about 5.8 sigma, 685 cells) stays inside the 2048-cell window
# centred on the bunch; the 4 um transverse cells put the CPML 11 sigma
# from the axis.
CAPACITY_C5 = Workload(
    name="capacity 256x256x2048 undulator, 6.4e7 e-",
    dims=(256, 256, 2048), spacing=(4.0e-6, 4.0e-6, 2.0e-6), cpml=16, particles=64_000_000,
    gamma_lab=500.0, gamma_boost=4.0, sigma=(40e-6, 40e-6, 30e-6), sigma_gamma=1e-4, emittance=1e-6,
    charge_electrons=64_000_000 * 355.0, dipole_tesla=0.0, dipole_length=0.0, seed=505, z_cells=1024.0,
    undulator=(1.0, 0.02, 6, 100, 0.5))


def undulator_line(K: float, period: float, segments: int, periods: int, gap: float, first_center: float = 0.0):
    """Planar undulator segments of strength K (peak field K 2 pi m_e c /
    (e lambda_u)) along the lab z axis: (center z, length, peak field)."""
    peak = K * 2.0 * np.pi * bunch.ME * C_SI / (bunch.QE * period)
    length = periods * period
    return [(first_center + s * (length + gap), length, peak) for s in range(segments)]


def quiet_start(electrons: float, length: float, wavelength: float, gamma: float, sigma_gamma: float,
                sigma_x: float, sigma_y: float, emittance_x: float, emittance_y: float, harmonics: int,
                beamlets: int, seed: int):
    """Multi-beamlet, multiharmonic quiet start of utilities/
    sase_bunch_generator.cpp (generate, noisyPhases, sampleTransverse,
    sampleGamma; lines 262-302, 432-463, 604-680): a top-hat of whole
    resonant-wavelength cells, `beamlets` Gaussian beamlets per cell, each
    cloned into 2 (harmonics + 1) members at quiet-start phases displaced by
    the shot-noise harmonics a_m cos(m phi) + b_m sin(m phi) (sigma =
    sqrt(2 / (m^2 N_e))). One splitmix64 stream (seed) feeds every draw in
    the generator's order; vectorised over beamlets, which is exact while
    every gamma draw is accepted at its first attempt (checked).

    Returns lab-frame positions relative to the bunch centre (n,3), proper
    velocities (n,3) and the one-based source ids, beamlet-major and
    member-major as the generator writes them."""
    cells = int(np.floor(length / wavelength + 0.5))
    if cells == 0:
        raise ValueError("Bunch length must contain at least one resonant wavelength")
    eff = cells * wavelength
    members = 2 * (harmonics + 1)
    nb = cells * beamlets
    per_beamlet = (electrons / cells) / beamlets
    draws = 2 * (4 + 1 + 2 * harmonics)  # uniforms per beamlet: 4 transverse, 1 gamma, 2 per harmonic normals
    idx = np.arange(nb, dtype=np.uint64)
    with np.errstate(over="ignore"):
        state0 = np.uint64(seed) + idx * np.uint64(draws) * bunch._GOLDEN

    def uniform(k):
        # the k-th draw of each beamlet's stream: state advanced k + 1 times
        with np.errstate(over="ignore"):
            z = state0 + np.uint64(k + 1) * bunch._GOLDEN
            z = (z ^ (z >> np.uint64(30))) * bunch._M1
            z = (z ^ (z >> np.uint64(27))) * bunch._M2
            z = z ^ (z >> np.uint64(31))
        return ((z >> np.uint64(11)).astype(np.float64) + 0.5) / 9007199254740992.0

    def normal(j):
        return np.sqrt(-2.0 * np.log(uniform(2 * j))) * np.cos(2.0 * np.pi * uniform(2 * j + 1))

    x1, x2, y1, y2 = normal(0), normal(1), normal(2), normal(3)
    x = sigma_x * x1
    y = sigma_y * y1
    ux = emittance_x / sigma_x * x2 if emittance_x else np.zeros(nb)
    uy = emittance_y / sigma_y * y2 if emittance_y else np.zeros(nb)
    g = gamma + sigma_gamma * normal(4)
    if np.any(g <= np.sqrt(1.0 + ux * ux + uy * uy)):
        raise ValueError("a gamma draw needs a second attempt; the vectorised stream no longer applies")
    uz = np.sqrt(g * g - 1.0 - ux * ux - uy * uy)
    quiet_order = harmonics + 1
    beamlet = (idx % np.uint64(beamlets)).astype(np.float64)
    cell = (idx // np.uint64(beamlets)).astype(np.float64)
    base = (beamlet + 0.5) * np.pi / (quiet_order * beamlets)
    disp = np.zeros((nb, members))
    quiet = base[:, None] + np.arange(members)[None, :] * np.pi / quiet_order
    for m in range(1, harmonics + 1):
        sig = np.sqrt(2.0 / (m * m * per_beamlet))
        a = sig * normal(5 + 2 * (m - 1))
        b = sig * normal(6 + 2 * (m - 1))
        disp += a[:, None] * np.cos(m * quiet) + b[:, None] * np.sin(m * quiet)
    phase = np.mod(quiet + disp, 2.0 * np.pi)
    n = nb * members
    pos = np.empty((n, 3))
    u = np.empty((n, 3))
    pos[:, 0] = np.repeat(x, members)
    pos[:, 1] = np.repeat(y, members)
    pos[:, 2] = (-0.5 * eff + (cell[:, None] + phase / (2.0 * np.pi)) * wavelength).ravel()
    u[:, 0] = np.repeat(ux, members)
    u[:, 1] = np.repeat(uy, members)
    u[:, 2] = np.repeat(uz, members)
    return pos, u, np.arange(1, n + 1, dtype=np.uint64)


@dataclass
class FelSegment:
    """Planar-undulator FEL segment: 4e6 macro-particles of weight 355,
    quiet-start loaded (8 phase-spaced clones per seeded 6D sample, shot
    noise through harmonic 3), lab gamma 1000, sigma_gamma/gamma 1e-4,
    K = 1, lambda_u = 2 cm, 6 segments of 100 periods, resonant boost
    gamma_b = gamma / sqrt(1 + K^2 / 2), 128 x 128 x 1024, quintic,
    FP32-split. The window holds a flat-top slice of the bunch whose box-frame
    length fills 70% of it (48 resonant wavelengths, 10417 beamlets each);
    the bunch is placed at the entrance of the first segment."""
    name: str = "FEL undulator segment 128x128x1024, 4e6 quiet-start e-"
    dims: tuple = (128, 128, 1024)
    spacing: tuple = (4.0e-6, 4.0e-6, 1.0e-6)
    cpml: int = 8
    gamma_lab: float = 1000.0
    sigma_gamma: float = 1e-4
    sigma: tuple = (30e-6, 30e-6)
    emittance: float = 0.5e-6
    weight: float = 355.0
    K: float = 1.0
    period: float = 0.02
    segments: int = 6
    periods: int = 100
    gap: float = 0.5
    harmonics: int = 3
    beamlets: int = 10417
    wavelength_cells: int = 48
    seed: int = 1
    z_cells: float = 512.0
    substeps: int = 1

    @property
    def gamma_boost(self) -> float:
        return self.gamma_lab / np.sqrt(1.0 + 0.5 * self.K ** 2)

    @property
    def dt(self) -> float:
        return self.spacing[2] / C_SI

    @property
    def cells(self) -> int:
        return int(np.prod(self.dims))

    @property
    def particles(self) -> int:
        return self.wavelength_cells * self.beamlets * 2 * (self.harmonics + 1)

    @property
    def resonant_wavelength(self) -> float:
        return self.period * (1.0 + 0.5 * self.K ** 2) / (2.0 * self.gamma_lab ** 2)

    def magnets(self):
        return undulator_line(self.K, self.period, self.segments, self.periods, self.gap,
                              first_center=0.5 * self.periods * self.period)

    def generate(self):
        """Box-frame particles: position (n,3), u (n,3), charge, mass, ids."""
        nx, ny, nz = self.dims
        dx, dy, dz = self.spacing
        lam = self.resonant_wavelength
        n = self.particles
        pos, u_lab, ids = quiet_start(n * self.weight, self.wavelength_cells * lam, lam, self.gamma_lab,
                                      self.sigma_gamma * self.gamma_lab, self.sigma[0], self.sigma[1],
                                      self.emittance, self.emittance, self.harmonics, self.beamlets, self.seed)
        u_mean = np.sqrt(self.gamma_lab ** 2 - 1.0)
        g_box = bunch.gamma_from_u(bunch.lab_to_box_u(np.array([[0.0, 0.0, u_mean]]), self.gamma_boost))[0]
        stretch = self.gamma_lab / g_box
        x = np.empty_like(pos)
        x[:, 0] = 0.5 * nx * dx + pos[:, 0]
        x[:, 1] = 0.5 * ny * dy + pos[:, 1]
        x[:, 2] = self.z_cells * dz + stretch * pos[:, 2]
        u = bunch.lab_to_box_u(u_lab, self.gamma_boost)
        q = np.full(n, -bunch.QE * self.weight)
        m = np.full(n, bunch.ME * self.weight)
        return x, u, q, m, ids


FEL_C3 = FelSegment()


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
    if getattr(w, "dipole_tesla", 0.0):
        coupling.addPrescribedLabMagnet("uniform-dipole", [0.0, 0.0, -0.5 * w.dipole_length], w.dipole_length,
                                        peak_field=w.dipole_tesla, polarization_angle=0.5 * np.pi)
    for zc, length, peak in magnets_of(w):
        coupling.addPrescribedLabMagnet("planar-undulator", [0.0, 0.0, zc], length, period=undulator_period(w),
                                        peak_field=peak)
    coupling.setParticleBoundary((w.cpml, w.cpml, 0))
    if data == "device":
        store = A.ParticleStore(fields, w.particles)
        w.generate_on_device(store)
        return fields, window, coupling, store
    x, u, q, m, ids = w.generate() if data is None else data
    store = A.ParticleStore(fields, len(q))
    store.upload(x, u, q, m, None, ids)
    return fields, window, coupling, store


def magnets_of(w):
    """Planar undulator segments (center z, length, peak field) of a workload."""
    if isinstance(w, FelSegment):
        return w.magnets()
    if getattr(w, "undulator", ()):
        K, period, segments, periods, gap = w.undulator
        return undulator_line(K, period, segments, periods, gap, first_center=0.5 * periods * period)
    return []


def undulator_period(w) -> float:
    return w.period if isinstance(w, FelSegment) else (w.undulator[1] if getattr(w, "undulator", ()) else 0.0)


class Stepper:
    """Drives Simulation::solve's per-step order (simulation.cpp:518-583):
    clear J, window begin, magnetic, coupled source step, shift."""

    def __init__(self, w: Workload, fields, window, coupling, store, offset: int = 10_000,
                 continuity: bool = False):
        self.w, self.fields, self.window, self.coupling, self.store = w, fields, window, coupling, store
        self.step_index = 0
        self.offset = offset
        self.origin = np.zeros(3)
        self.time = 0.0
        self.trials = []
        # (max |(rho1 - rho0)/dt + div J|, max |(rho1 - rho0)/dt|,
        # max |rho|/dt) of every step, read before the shift clears J
        self.track_continuity = continuity
        self.continuity = []
        # start each step's sort beside the field updates (aprl_coupling_prepare_step)
        self.prepare = True

    def step(self, shift: bool = True):
        """One step; shift=False stops after the coupled source step (J of
        the step still on the grid, the window not yet moved) and leaves
        the stepper there."""
        f, win = self.fields, self.window
        n = self.step_index
        # the sort phase runs on its own stream beside the field updates
        if self.prepare:
            self.coupling.prepareStep(self.store, self.origin)
        f.clearCurrent()
        win.beginStep(n, self.offset)
        f.advanceMagnetic()
        win.setInlet()
        its, _, _ = self.coupling.step(self.store, self.origin, self.time, getattr(self.w, "substeps", 1))
        self.trials.append(its)
        if self.track_continuity:
            self.continuity.append(self.coupling.continuity(self.store))
        if not shift:
            return
        self.offset -= 1
        win.shift(n + 1, self.offset)
        # Simulation::finishMovingWindowStep: origin = fma(offset, dz, ref)
        # with the reference origin at offset 10000 (fma(a, b, 0) = a*b).
        self.origin[2] = (self.offset - 10_000) * self.w.spacing[2]
        self.time += self.w.dt
        self.step_index += 1


@dataclass
class InjectionSetup:
    """Planes and window placement of a frozen lab-plane injection run: the
    input (= activation) plane, the initial bunch centre and the first
    magnetic interaction entrance in the lab (m); the box z (cells) of the
    window's lower face at step 0, and of the lab exit plane and the lab
    field plane at box time zero."""
    input_plane_z: float
    initial_center_z: float
    entrance: float | None
    window_bottom_cells: float
    exit_box_cells: float
    field_plane_box_cells: float


# The bending-radiation card as a production run: the bunch's plane-arrival
# events 1 mm upstream of the dipole entrance (z_lab = -0.1 m), projected to a
# snapshot centred 0.1 mm further on, frozen and injected through the
# pending queue; released particles cross the window (about two cells per
# step relative to it) and leave through the lab exit plane, which the
# window passes at 0.02 cells per step (it stays 32 cells below the upper
# face over 2000 steps); the lab field plane sits inside the window for the
# whole run.
BENDING_INJECTION = InjectionSetup(input_plane_z=-0.101, initial_center_z=-0.1009, entrance=-0.1,
                                   window_bottom_cells=-16.0, exit_box_cells=440.0, field_plane_box_cells=200.0)


class ProductionRun:
    """Simulation's moving-window run for workload w from plane-arrival
    events (simulation.cpp:1983-2245, then solve's per-step order,
    simulation.cpp:518-597): the events of the seeded Gaussian are drawn on
    the device, pass the input route (aprl_amd.injection), the snapshot's
    upstream particles wait in the library's pending queue and the rest
    enter the store held; every step admits the due particles (8-step
    look-ahead), runs the field half step, the coupled source step and the
    energy-ledger flux, shifts the window and samples the lab field plane
    at every step's lab time."""

    def __init__(self, w: Workload, setup: InjectionSetup, shape: str = "linear-binomial3",
                 precision: str = "fp64", max_steps: int = 2000, plane_capacity: int = 64, keep_patch=None):
        import aprl_amd as A
        from . import pending
        from .injection import InjectionPlanes, LabPlaneBunch, frozenLabPlaneInjection
        self.w, self.setup = w, setup
        nx, ny, nz = w.dims
        dx, dy, dz = w.spacing
        g = A.EBGridGeometry(*w.dims, *w.spacing, w.dt)
        self.fields = A.EBFieldGrid(g, precision=precision)
        self.window = A.MovingWindowFields(self.fields, A.EBCPMLParameters((w.cpml, w.cpml, 0)))
        self.coupling = A.ConservativeSourceCoupling(self.fields, shape)
        if w.dipole_tesla:
            self.coupling.addPrescribedLabMagnet("uniform-dipole", [0.0, 0.0, -0.5 * w.dipole_length],
                                                 w.dipole_length, peak_field=w.dipole_tesla,
                                                 polarization_angle=0.5 * np.pi)
        # the plane-arrival events and the input route
        self.bunch = LabPlaneBunch()
        u_lab = np.sqrt(w.gamma_lab ** 2 - 1.0)
        beta_lab = u_lab / w.gamma_lab
        sig_u = (w.emittance / w.sigma[0], w.emittance / w.sigma[1], w.sigma_gamma * w.gamma_lab)
        self.bunch.generatePlaneEvents(w.particles, (0.0, 0.0), w.sigma[:2], w.sigma[2] / (beta_lab * C_SI),
                                       (0.0, 0.0, u_lab), sig_u, w.charge_electrons, w.seed)
        planes = InjectionPlanes(setup.input_plane_z, setup.input_plane_z, setup.initial_center_z, setup.entrance)
        self.frame, self.reports = frozenLabPlaneInjection(self.bunch, planes, w.gamma_boost, dz)
        x, u, q, m, wt, ids, beta = self.bunch.download()
        self.coupling.setFrame(*self.frame)
        self.exit_lab = self.frame[2] + w.gamma_boost * setup.exit_box_cells * dz
        self.coupling.setParticleBoundary((w.cpml, w.cpml, 0), self.exit_lab)
        self.coupling.setActivationPlane(setup.input_plane_z)
        self.coupling.setHoldTransverse(True)
        # the window: lower face at window_bottom_cells at step 0
        self.offset = 10_000
        ref = setup.window_bottom_cells * dz - self.offset * dz
        self.queue = pending.PendingParticleQueue(
            pending.PendingWindowGeometry(ref, self.offset, dz, C_SI * w.dt, max_steps + 64))
        xs, us, qs, ms, ws, idss, fs, bs = self.queue.extract(x, u, q, m, wt, ids, np.ones(len(q), np.int32), beta)
        self.store = A.ParticleStore(self.fields, len(q))
        self.store.upload(xs, us, qs, ms, ws, idss)
        self.store.setHeld(fs, bs)
        self.queued0 = self.queue.size()
        self.origin = np.array([-0.5 * nx * dx, -0.5 * ny * dy, self.queue.origin(0)])
        self.window.enableEnergyLedger(True)
        # the lab field plane, sampled once per step (the lab time of a fixed
        # lab z advances dt / gamma_b per step)
        z_plane = self.frame[2] + w.gamma_boost * setup.field_plane_box_cells * dz
        self.plane = A.LabFieldPlane(self.coupling, z_plane, 0.999 * w.dt / w.gamma_boost, plane_capacity)
        self.plane_samples = 0
        # (i0, i1, j0, j1): the part of every fetched plane record kept on the host
        self.keep_patch = keep_patch
        self.kept_times, self.kept = [], []
        self.step_index = 0
        self.time = 0.0
        self.trials = []
        self.particle_steps = 0

    def step(self):
        n = self.step_index
        self.queue.admit(n, self.store, 8)
        self.particle_steps += len(self.store)
        f, win = self.fields, self.window
        f.clearCurrent()
        win.beginStep(n, self.offset)
        f.advanceMagnetic()
        win.setInlet()
        its, _, _ = self.coupling.step(self.store, self.origin, self.time, self.w.substeps)
        self.trials.append(its)
        win.finishEnergyLedgerStep()
        self.offset -= 1
        win.shift(n + 1, self.offset)
        self.step_index += 1
        self.time += self.w.dt
        self.origin[2] = self.queue.origin(self.step_index)
        if self.plane.sample(self.time, self.origin):
            self.plane_samples += 1
            if self.plane_samples % self.plane.capacity == 0:
                self._fetch()

    def _fetch(self):
        times, records = self.plane.fetch()
        if self.keep_patch is not None and len(times):
            i0, i1, j0, j1 = self.keep_patch
            self.kept_times.append(times)
            self.kept.append(records[:, j0:j1, i0:i1, :].copy())

    def finish(self):
        """Fetches the plane samples still in the device ring; returns the
        kept (lab times, records) of the whole run."""
        self._fetch()
        if not self.kept:
            return np.zeros(0), np.zeros((0,))
        return np.concatenate(self.kept_times), np.concatenate(self.kept)


class BatchStepper:
    """Steps several independent instances (each a workload with its own
    grid, window, coupling and store) with aprl_window_step_batch: the
    per-instance bookkeeping of Stepper, one C-ABI call per step."""

    def __init__(self, instances):
        import ctypes as C
        from . import _capi
        self.instances = instances  # (workload, fields, window, coupling, store)
        self.steps = (_capi.InstanceStep * len(instances))()
        for s, (w, f, win, cg, st) in zip(self.steps, instances):
            s.fields, s.coupling, s.particles = f.handle, cg.handle, st.handle
            # the window's offset counts from 0 at the start, with the
            # reference origin at 0: origin z = fma(offset, dz, 0)
            s.step, s.offset, s.time, s.substeps = 0, 0, 0.0, 1
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
            s.origin[2] = s.offset * w.spacing[2]
            s.time += w.dt

    def run(self, steps: int):
        """`steps` steps with one aprl_window_run_batch call (the library
        advances step, offset, origin and time; an instance's next step is
        enqueued as soon as its own is committed). Returns the Picard trials
        of each instance summed over the steps."""
        from . import _capi
        from ._capi import check, lib
        runs = (_capi.InstanceRun * len(self.instances))()
        for r, s in zip(runs, self.steps):
            r.state = s
            r.reference_z = 0.0
        check(lib().aprl_window_run_batch(runs, len(self.instances), int(steps)))
        for r, s in zip(runs, self.steps):
            self._C.memmove(self._C.addressof(s), self._C.addressof(r.state), self._C.sizeof(s))
        return [int(r.iterations) for r in runs]


def batch_injection(w: Workload) -> InjectionSetup:
    """The bending card's planes for a batched-scan instance: events 1 mm
    upstream of its dipole, the window's lower face 16 cells below the box
    origin, the lab exit plane 72 cells below the upper face and the lab
    field plane at 40% of the window at box time zero."""
    nz = w.dims[2]
    return InjectionSetup(input_plane_z=-0.101, initial_center_z=-0.1009, entrance=-0.1, window_bottom_cells=-16.0,
                          exit_box_cells=float(nz - 72), field_plane_box_cells=0.4 * nz)


class BatchProduction:
    """Several production runs (ProductionRun, each its own card) advanced
    together by aprl_window_run_batch: in every step each instance admits
    its due particles, takes the window step and samples its lab field
    plane, and an instance's next step is enqueued as soon as its own is
    committed. Runs are cut at the plane rings' capacity, whose samples are
    then fetched (the fetch waits for the device)."""

    def __init__(self, runs):
        from . import _capi
        self.runs = runs
        self.c = (_capi.InstanceRun * len(runs))()
        for r, run in zip(self.c, runs):
            s = r.state
            s.fields, s.coupling, s.particles = run.fields.handle, run.coupling.handle, run.store.handle
            s.step, s.offset, s.time, s.substeps = run.step_index, run.offset, run.time, run.w.substeps
            for a in range(3):
                s.origin[a] = run.origin[a]
            r.queue = run.queue.handle
            r.look_ahead = 8
            r.plane = run.plane.handle

    def run(self, steps: int):
        from ._capi import check, lib
        done = 0
        cap = min(run.plane.capacity for run in self.runs)
        while done < steps:
            k = min(steps - done, cap)
            check(lib().aprl_window_run_batch(self.c, len(self.runs), k))
            for r, run in zip(self.c, self.runs):
                run.plane_samples = int(r.samples)
                run._fetch()
            done += k
        for r, run in zip(self.c, self.runs):
            s = r.state
            run.step_index, run.offset, run.time = int(s.step), int(s.offset), float(s.time)
            run.origin[:] = [s.origin[0], s.origin[1], s.origin[2]]
            run.particle_steps = int(r.entry_steps)

    @property
    def iterations(self):
        return [int(r.iterations) for r in self.c]


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
