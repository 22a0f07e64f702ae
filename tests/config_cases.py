"""The benchmark configurations driven through the oracle beside the GPU
(test infrastructure: the product never imports it).

``OracleStepper`` runs Simulation::solve's per-step order (simulation.cpp:
518-583) on the oracle with the zero inlet reference, the same step, offset
and origin sequence as ``aprl_amd.workloads.Stepper``.
"""
from __future__ import annotations

import numpy as np

import harness as H
from aprl_amd import workloads as W

SHAPES = {"linear-binomial3": ("binomial3", True), "quintic-compensated": ("quintic", False)}


def oracle_for(w, shape: str, data):
    """Oracle field grid + window + coupling of workload w with the
    particles `data` (x, u, q, m, ids)."""
    oshape, smooth = SHAPES[shape]
    orc = H.CpuFields("orc", *w.dims, *w.spacing, w.dt, 1)
    orc.install_window((w.cpml, w.cpml, 0))
    co = H.CpuCoupling(orc, oshape, smooth, cpml_cells=(w.cpml, w.cpml, 0))
    co.set_frame(w.gamma_boost)
    if getattr(w, "dipole_tesla", 0.0):
        co.add_magnet("dipole", [0.0, 0.0, -0.5 * w.dipole_length], w.dipole_length, peak=w.dipole_tesla,
                      angle=0.5 * np.pi)
    for zc, length, peak in W.magnets_of(w):
        co.add_magnet("undulator", [0.0, 0.0, zc], length, period=W.undulator_period(w), peak=peak)
    x, u, q, m, ids = data
    co.set_particles(x, u, q, m, np.ones(len(q)), ids)
    return orc, co


class OracleStepper:
    def __init__(self, w, orc, co, offset: int = 10_000):
        self.w, self.orc, self.co = w, orc, co
        self.offset = offset
        self.origin = np.zeros(3)
        self.time = 0.0
        self.n = 0
        self.trials = []
        self.zero = [np.zeros(int(np.prod(H.component_shape(*w.dims, a)[1:]))) for a in range(6)]

    def step(self, shift: bool = True):
        o, z = self.orc, self.zero
        o.clear_current()
        o.window_begin(self.n, self.offset, z[:3])
        o.window_advance_magnetic()
        o.window_set_inlet(self.offset, self.n + 1, z[0], z[1])
        its, _, _ = self.co.step(self.origin, self.time, getattr(self.w, "substeps", 1))
        self.trials.append(its)
        if not shift:
            return
        self.offset -= 1
        o.window_shift(self.n + 1, self.offset, z)
        self.origin[2] = (self.offset - 10_000) * self.w.spacing[2]
        self.time += self.w.dt
        self.n += 1


def rel_l2(a, b):
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def field_errors(gpu, orc):
    return {a: rel_l2(gpu.component(a), orc.view(a)) for a in range(9)}


def vector_errors(gpu, orc):
    """Relative L2 error of E, B and J as vector fields: |test - ref| over
    the three components together, divided by |ref| of the three (a
    component that is small beside the others, such as Bz of a
    relativistic bunch, cancels in the curl that forms it and carries
    the absolute round-off of the larger ones)."""
    out = {}
    for name, comps in (("E", (0, 1, 2)), ("B", (3, 4, 5)), ("J", (6, 7, 8))):
        num = sum(float(np.sum((gpu.component(a) - orc.view(a)) ** 2)) for a in comps)
        den = sum(float(np.sum(orc.view(a) ** 2)) for a in comps)
        out[name] = float(np.sqrt(num / den)) if den > 0 else float(np.sqrt(num))
    return out


def continuity_ok(record, tol):
    """A step's continuity record (residual, rate, rho scale) within tol
    of the charge-density scale max |rho|/dt: the residual's round-off is
    that of rho and div J, each summed from many taps."""
    residual, rate, scale = record
    return rate > 0 and residual <= tol * scale


def particle_errors(store, co):
    xo, uo = co.particles()
    xg, ug, _, _, _, ids = store.download()
    order = np.argsort(ids)
    ido = co.ids()
    assert np.array_equal(np.sort(ids), np.sort(ido)), "the GPU store and the oracle hold different ids"
    oo = np.argsort(ido)
    xo, uo = xo[oo], uo[oo]
    return rel_l2(xg[order] - xo.mean(0), xo - xo.mean(0)), rel_l2(ug[order], uo)
