"""Host-side mirror of the particle and coupling interfaces over the C-ABI.

``ParticleStore`` holds RelativisticParticleSI records on the device
(eb_particles.h:12-25); ``ConservativeSourceCoupling`` keeps the name of
source_coupling.h:12-41 and runs one whole
Simulation::solveConservativeSourceStep + commitConservativeElectric
(simulation_source_coupling.cpp:16-65) per ``step`` call.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _capi
from ._capi import check, dptr, lib
from .fields import EBFieldGrid

SHAPES = {"quintic-compensated": 0, "linear-binomial3": 2}
MAGNETS = {"planar-undulator": 0, "uniform-dipole": 1, "quadrupole": 2}


class ParticleStore:
    def __init__(self, fields: EBFieldGrid, capacity: int = 1):
        self.fields = fields
        h = C.c_void_p()
        check(lib().aprl_particles_create(fields.handle, capacity, C.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().aprl_particles_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return lib().aprl_particles_count(self.handle)

    def upload(self, position, u, charge, mass, weight=None, ids=None):
        x = np.ascontiguousarray(position, dtype=np.float64).reshape(-1, 3)
        n = x.shape[0]
        uu = np.ascontiguousarray(u, dtype=np.float64).reshape(n, 3)
        q = np.ascontiguousarray(np.broadcast_to(charge, (n,)), dtype=np.float64)
        m = np.ascontiguousarray(np.broadcast_to(mass, (n,)), dtype=np.float64)
        w = np.ascontiguousarray(np.ones(n) if weight is None else np.broadcast_to(weight, (n,)), dtype=np.float64)
        i = np.ascontiguousarray(np.arange(1, n + 1) if ids is None else ids, dtype=np.uint64)
        check(lib().aprl_particles_upload(self.handle, n, dptr(x.ravel()), dptr(uu.ravel()), dptr(q), dptr(m),
                                          dptr(w), i.ctypes.data_as(C.POINTER(C.c_uint64))))

    def generateGaussian(self, count, center, sigma_position, mean_u, sigma_u, electrons, seed, first=0, total=None,
                         boost_gamma=1.0):
        """ParticleInitializer's generated-gaussian beam drawn on the device
        (aprl_particles_generate_gaussian), replacing the store's contents."""
        from ._capi import GaussianBeam
        b = GaussianBeam()
        for a in range(3):
            b.center[a], b.sigma_position[a] = center[a], sigma_position[a]
            b.mean_u[a], b.sigma_u[a] = mean_u[a], sigma_u[a]
        b.electrons, b.seed, b.first = float(electrons), int(seed), int(first)
        b.total = int(count if total is None else total)
        b.boost_gamma = float(boost_gamma)
        check(lib().aprl_particles_generate_gaussian(self.handle, C.byref(b), int(count)))

    def download(self):
        """(position (n,3), u (n,3), charge, mass, weight, ids) in device order."""
        n = len(self)
        x = np.empty((n, 3))
        u = np.empty((n, 3))
        q, m, w = np.empty(n), np.empty(n), np.empty(n)
        ids = np.empty(n, dtype=np.uint64)
        check(lib().aprl_particles_download(self.handle, dptr(x.ravel()) if n else None,
                                            dptr(u.ravel()) if n else None, dptr(q) if n else None,
                                            dptr(m) if n else None, dptr(w) if n else None,
                                            ids.ctypes.data_as(C.POINTER(C.c_uint64)) if n else None))
        return x, u, q, m, w, ids

    def setHeld(self, frozen, beta):
        """RelativisticParticleSI::frozen / frozenBetaZ for the entries in
        upload order (before the first step)."""
        fr = np.ascontiguousarray(frozen, dtype=np.int32)
        b = np.ascontiguousarray(beta, dtype=np.float64)
        check(lib().aprl_particles_set_held(self.handle, fr.ctypes.data_as(C.POINTER(C.c_int)), dptr(b)))

    def states(self):
        """Entry state per live entry in download order: 0 particle, 1 CPML
        carrier (ParticleCPMLCarrier, particle_boundary.h:76-84)."""
        n = len(self)
        st = np.zeros(n, dtype=np.int32)
        if n:
            check(lib().aprl_particles_download_states(self.handle, st.ctypes.data_as(C.POINTER(C.c_int))))
        return st


class ConservativeSourceCoupling:
    def __init__(self, fields: EBFieldGrid, shape: str = "linear-binomial3", transverseXBinomial3: bool = True):
        self.fields = fields
        h = C.c_void_p()
        check(lib().aprl_coupling_create(fields.handle, SHAPES[shape], int(transverseXBinomial3), C.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().aprl_coupling_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setFrame(self, gamma, labTimeOrigin=0.0, labZOrigin=0.0, boxZOrigin=0.0):
        b = _capi.Boost(gamma, labTimeOrigin, labZOrigin, boxZOrigin)
        check(lib().aprl_coupling_set_frame(self.handle, C.byref(b)))

    def addPrescribedLabMagnet(self, kind, center, length, period=0.0, peak_field=0.0, polarization_angle=0.0,
                               gaussian_fringe=False, fringe_relative_cutoff=1e-9, quadrupole_gradient=0.0,
                               end_ramp_periods=0.0, undulator_phase=0.0):
        m = _capi.Magnet()
        m.type = MAGNETS[kind]
        for a in range(3):
            m.center[a] = center[a]
        m.length, m.period, m.peak_field, m.polarization_angle = length, period, peak_field, polarization_angle
        m.gaussian_fringe = int(gaussian_fringe)
        m.fringe_relative_cutoff = fringe_relative_cutoff
        m.quadrupole_gradient = quadrupole_gradient
        m.end_ramp_periods = end_ramp_periods
        m.undulator_phase = undulator_phase
        check(lib().aprl_coupling_add_magnet(self.handle, C.byref(m)))

    def setParticleBoundary(self, cpml_cells, exit_lab_z=math.inf):
        cells = (C.c_size_t * 3)(*cpml_cells)
        check(lib().aprl_coupling_set_particle_boundary(self.handle, cells, exit_lab_z))

    def setReferenceCurrent(self, table=None, guards=(0, 0), translation_beta=0.0):
        """TransverseReferenceCurrent with its prepared table (rows, points);
        None disables it."""
        if table is None:
            check(lib().aprl_coupling_set_reference_current(self.handle, None, 0, 0, None, 0.0))
            return
        t = np.ascontiguousarray(table, dtype=np.float64)
        g = (C.c_size_t * 2)(*[int(v) for v in guards])
        check(lib().aprl_coupling_set_reference_current(self.handle, dptr(t.ravel()), t.shape[0], t.shape[1], g,
                                                        float(translation_beta)))

    def setActivationPlane(self, lab_z):
        """Lab z of the held particles' release (NaN disables it)."""
        check(lib().aprl_coupling_set_activation(self.handle, float(lab_z)))

    def setHoldTransverse(self, hold: bool):
        """config reference.hold_transverse_until_activation: held particles
        move z += c dt beta (True, the default here) or ballistically along
        their proper velocity (False, the reference's default)."""
        check(lib().aprl_coupling_set_hold_transverse(self.handle, int(bool(hold))))

    def boundaryCounters(self):
        """Cumulative (counts[20], charges[20]): CPML entries per face (0-5),
        outer exits of particles (6-11) and of carriers (12-17), window exits
        (18) and cropped carriers (19)."""
        counts = np.zeros(20, dtype=np.uint64)
        charges = np.zeros(20)
        check(lib().aprl_coupling_boundary_counters(self.handle, counts.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                    dptr(charges)))
        return counts, charges

    def setSplitTrials(self, trials=2, verify=False):
        """FP32-split grids: run exactly ``trials`` Picard trials per step in
        place of the FP64 convergence test; ``verify`` adds a residual gather
        after the last trial whose force and work the step reports."""
        check(lib().aprl_coupling_set_split_trials(self.handle, int(trials), 1 if verify else 0))

    def setTrialBatch(self, batch=0):
        """Plain FP64 grids: Picard trials enqueued before the first wait for
        the device's convergence decision (0: the last step's count)."""
        check(lib().aprl_coupling_set_trial_batch(self.handle, int(batch)))

    def prepareStep(self, particles: ParticleStore, origin):
        """Start the next step's sort phase on the coupling's own stream
        (aprl_coupling_prepare_step); call before the step's field updates."""
        o = np.ascontiguousarray(origin, dtype=np.float64)
        check(lib().aprl_coupling_prepare_step(self.handle, particles.handle, dptr(o)))

    def step(self, particles: ParticleStore, origin, time, substeps=1):
        o = np.ascontiguousarray(origin, dtype=np.float64)
        r = _capi.StepReport()
        check(lib().aprl_coupling_step(self.handle, particles.handle, dptr(o), time, substeps, C.byref(r)))
        self.last_tiles, self.last_chunks, self.last_active_tiles = r.tiles, r.chunks, r.active_tiles
        return r.iterations, r.force, r.work

    def particleEnergy(self, particles: ParticleStore, time):
        """Energy-ledger particle sums (simulation.cpp:3591-3625): kinetic,
        total, represented electrons, represented-weighted lab gamma, box
        gamma, lab z, lab kinetic energy, c p_z box, c p_z lab."""
        out = np.zeros(9)
        check(lib().aprl_coupling_particle_energy(self.handle, particles.handle, float(time), dptr(out)))
        return out

    def continuity(self, particles: ParticleStore):
        """(max |(rho1 - rho0)/dt + div J|, max |(rho1 - rho0)/dt|,
        max max(|rho0|, |rho1|)/dt) of the last step
        (ChargeConservingCurrentDepositor::maxContinuityResidual,
        eb_deposition.cpp:195-226); call before the window shift."""
        out = np.zeros(3)
        check(lib().aprl_coupling_continuity(self.handle, particles.handle, dptr(out)))
        return out[0], out[1], out[2]

    def gaussResidual(self, particles: ParticleStore, origin, lo, hi):
        """(sum err^2, sum src^2, max |err|, max |src|) of div E - rho/eps0
        over the vertices strictly inside [lo, hi] (eb_initial_field.cpp:
        774-799)."""
        out = np.zeros(4)
        o = np.ascontiguousarray(origin, dtype=np.float64)
        check(lib().aprl_coupling_gauss(self.handle, particles.handle, dptr(o), _capi.size3(lo), _capi.size3(hi),
                                        dptr(out)))
        return tuple(out)


class LabFieldPlane:
    """One laboratory field plane (FieldDetectorPlaneConfig: lab z and lab-time
    rhythm) of LabDetectorManager::sampleFieldPlanes (lab_detectors.cpp:
    2426-2487), sampled on the device with the coupling's frame and grid."""

    def __init__(self, coupling: ConservativeSourceCoupling, z_lab: float, rhythm: float, capacity: int = 64):
        self.coupling = coupling
        g = coupling.fields.geometry
        self.points = g.nx * g.ny
        self.shape = (g.ny, g.nx, 6)
        self.capacity = capacity
        h = C.c_void_p()
        check(lib().aprl_lab_plane_create(coupling.handle, float(z_lab), float(rhythm), capacity, C.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().aprl_lab_plane_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def sample(self, time_box: float, origin) -> bool:
        """sampleFieldPlanes at the step's end box time and shifted origin."""
        o = np.ascontiguousarray(origin, dtype=np.float64)
        taken = C.c_int(0)
        check(lib().aprl_lab_plane_sample(self.handle, float(time_box), dptr(o), C.byref(taken)))
        return bool(taken.value)

    def fetch(self):
        """(lab times (n,), records (n, ny, nx, 6): E_lab then B_lab) of the
        samples taken since the last fetch."""
        times = np.zeros(self.capacity)
        records = np.zeros((self.capacity, *self.shape))
        n = _capi._sz(0)
        check(lib().aprl_lab_plane_fetch(self.handle, self.capacity, dptr(times), dptr(records.reshape(-1)),
                                         C.byref(n)))
        return times[:n.value].copy(), records[:n.value].copy()


class EBGaussFieldInitializer:
    """EBGaussFieldInitializer (eb_initial_field.h:37-70) on the device: the
    Gauss-consistent initial E/B of the store's particles written into the
    grid (aprl_fields_initialize_gauss)."""

    MODELS = {"relativistic-poisson": 0, "electrostatic-poisson": 1, "cowan-relativistic-poisson": 2}

    @staticmethod
    def initialize(fields, particles: ParticleStore, origin, model: str = "relativistic-poisson",
                   maximumMeanTransverseBeta: float = 1e-8, staticFieldGuardCells=(0, 0, 0),
                   relativeTolerance: float = 1e-10, maximumIterations: int = 10000,
                   includeTransverseCpmlParticles: bool = False) -> dict:
        o = np.ascontiguousarray(origin, dtype=np.float64)
        r = _capi.GaussReport()
        check(lib().aprl_fields_initialize_gauss(fields.handle, particles.handle, dptr(o),
                                                 EBGaussFieldInitializer.MODELS[model],
                                                 float(maximumMeanTransverseBeta), _capi.size3(staticFieldGuardCells),
                                                 float(relativeTolerance), int(maximumIterations),
                                                 int(includeTransverseCpmlParticles), C.byref(r)))
        return {"iterations": r.iterations, "relativeResidual": r.relative_residual,
                "maximumGaussResidual": r.maximum_gauss_residual, "totalCharge": r.total_charge,
                "meanBetaZ": r.mean_beta_z, "meanBetaTransverse": r.mean_beta_transverse,
                "initialElectricEnergy": r.initial_electric_energy, "initialMagneticEnergy": r.initial_magnetic_energy,
                "rmsPositionCells": list(r.rms_position_cells), "centrePaddingRms": list(r.centre_padding_rms),
                "staticFieldGuardCells": list(r.static_field_guard_cells),
                "transversePoissonPaddingCells": r.transverse_poisson_padding_cells}
