"""ctypes harness over the oracle (oracle/liboracle.so) and the reference build
(oracle/_ref/libaprl_ref.so). Both export the same entry points with the
prefixes ``orc_`` and ``ref_``; :class:`CpuFields` drives either one.

This module is test infrastructure: the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
REF_LIB = os.path.join(ORACLE_DIR, "_ref", "libaprl_ref.so")
REFERENCE_TREE = "/root/reference"

C_SI = 299792458.0
EPS0 = 8.8541878128e-12
MU0 = 1.25663706212e-6

_dp = C.POINTER(C.c_double)
_libs: dict[str, C.CDLL] = {}


def build_oracle() -> None:
    """Compile liboracle.so (and _ref when the reference tree is present),
    from scratch on every call."""
    targets = ["liboracle.so"]
    if os.path.isdir(REFERENCE_TREE):
        targets.append("ref")
    subprocess.run(["make", "-s", "-B", "-C", ORACLE_DIR, *targets], check=True)


def have_ref() -> bool:
    return os.path.exists(REF_LIB)


def load(prefix: str) -> C.CDLL:
    if prefix in _libs:
        return _libs[prefix]
    path = ORACLE_LIB if prefix == "orc" else REF_LIB
    if not os.path.exists(path):
        build_oracle()
    lib = C.CDLL(path)
    P = prefix + "_"
    sz = C.c_size_t
    getattr(lib, P + "last_error").restype = C.c_char_p
    getattr(lib, P + "cowan").argtypes = [C.c_double] * 3 + [_dp]
    getattr(lib, P + "fields_new").argtypes = [sz, sz, sz, C.c_double, C.c_double, C.c_double, C.c_double,
                                               C.c_int, C.POINTER(C.c_void_p)]
    getattr(lib, P + "fields_free").argtypes = [C.c_void_p]
    getattr(lib, P + "fields_free").restype = None
    getattr(lib, P + "fields_component").argtypes = [C.c_void_p, C.c_int, C.POINTER(sz)]
    getattr(lib, P + "fields_component").restype = _dp
    getattr(lib, P + "fields_history").argtypes = [C.c_void_p, C.c_int, C.POINTER(sz)]
    getattr(lib, P + "fields_history").restype = _dp
    getattr(lib, P + "fields_install_cpml").argtypes = [C.c_void_p, C.POINTER(sz), C.c_double, C.c_double,
                                                        C.c_double, C.c_double, sz, sz, C.c_int, C.c_int]
    getattr(lib, P + "fields_install_window").argtypes = [C.c_void_p, C.POINTER(sz), C.c_double, C.c_double,
                                                          C.c_double, C.c_double]
    getattr(lib, P + "fields_advance_magnetic").argtypes = [C.c_void_p, C.POINTER(_dp)]
    for name in ("fields_advance_electric", "fields_clear_current", "window_advance_magnetic",
                 "window_undo_preview", "window_commit"):
        getattr(lib, P + name).argtypes = [C.c_void_p]
    getattr(lib, P + "window_preview").argtypes = [C.c_void_p, C.c_int]
    getattr(lib, P + "window_begin").argtypes = [C.c_void_p, sz, C.c_ssize_t, C.POINTER(_dp)]
    getattr(lib, P + "window_set_inlet_at").argtypes = [C.c_void_p, C.c_ssize_t, sz, _dp, _dp]
    getattr(lib, P + "window_shift").argtypes = [C.c_void_p, sz, C.c_ssize_t, C.POINTER(_dp)]
    for name in ("fields_energy_density",):
        getattr(lib, P + name).argtypes = [C.c_void_p, sz, sz, sz]
        getattr(lib, P + name).restype = C.c_double
    getattr(lib, P + "fields_poynting").argtypes = [C.c_void_p, sz, sz, sz, C.c_uint]
    getattr(lib, P + "fields_poynting").restype = C.c_double
    _libs[prefix] = lib
    return lib


class CpuError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def component_shape(nx, ny, nz, a):
    """(nz, ny, nx) numpy shape of Yee component a (0..8: E, B, J)."""
    electric = a < 3 or a >= 6
    axis = a % 3
    cells = [nx, ny, nz]
    s = [cells[d] + ((d != axis) if electric else (d == axis)) for d in range(3)]
    return (s[2], s[1], s[0])


def _ptr(arr: np.ndarray):
    assert arr.dtype == np.float64 and arr.flags["C_CONTIGUOUS"]
    return arr.ctypes.data_as(_dp)


def _ptr_array(arrays):
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in arrays]
    arr = (_dp * len(keep))(*[_ptr(a) for a in keep])
    return arr, keep


class CpuFields:
    """One field grid in the oracle (``prefix='orc'``) or reference build."""

    def __init__(self, prefix, nx, ny, nz, dx, dy, dz, dt, solver=1):
        self.lib = load(prefix)
        self.P = prefix + "_"
        self.nx, self.ny, self.nz = nx, ny, nz
        h = C.c_void_p()
        self._check(self._fn("fields_new")(nx, ny, nz, dx, dy, dz, dt, solver, C.byref(h)))
        self.h = h

    def _fn(self, name):
        return getattr(self.lib, self.P + name)

    def _check(self, code):
        if code:
            raise CpuError(code, self._fn("last_error")().decode())

    def close(self):
        if getattr(self, "h", None):
            self._fn("fields_free")(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def view(self, a) -> np.ndarray:
        n = C.c_size_t()
        p = self._fn("fields_component")(self.h, a, C.byref(n))
        arr = np.ctypeslib.as_array(p, shape=(n.value,))
        return arr.reshape(component_shape(self.nx, self.ny, self.nz, a))

    def history(self, index) -> np.ndarray:
        n = C.c_size_t()
        p = self._fn("fields_history")(self.h, index, C.byref(n))
        if n.value == 0:
            return np.zeros(0)
        return np.ctypeslib.as_array(p, shape=(n.value,))

    def install_cpml(self, cells, order=3.0, R=1e-8, kappa=8.0, alpha=0.0, global_nz=None, zoff=0,
                     lower_z=True, upper_z=True):
        c = (C.c_size_t * 3)(*cells)
        self._check(self._fn("fields_install_cpml")(self.h, c, order, R, kappa, alpha,
                                                    global_nz or self.nz, zoff, int(lower_z), int(upper_z)))

    def install_window(self, cells, order=3.0, R=1e-8, kappa=8.0, alpha=0.0):
        c = (C.c_size_t * 3)(*cells)
        self._check(self._fn("fields_install_window")(self.h, c, order, R, kappa, alpha))

    def advance_magnetic(self, halo=None):
        if halo is None:
            self._check(self._fn("fields_advance_magnetic")(self.h, None))
        else:
            planes = [np.zeros(1) if p is None else p for p in halo]
            arr, keep = _ptr_array(planes)
            for i, p in enumerate(halo):
                if p is None:
                    arr[i] = _dp()
            self._check(self._fn("fields_advance_magnetic")(self.h, arr))

    def advance_electric(self):
        self._check(self._fn("fields_advance_electric")(self.h))

    def clear_current(self):
        self._check(self._fn("fields_clear_current")(self.h))

    # Moving-window protocol.
    def window_begin(self, step, offset, lower3):
        arr, keep = _ptr_array(lower3)
        self._check(self._fn("window_begin")(self.h, step, offset, arr))

    def window_advance_magnetic(self):
        self._check(self._fn("window_advance_magnetic")(self.h))

    def window_begin_centered(self, step, offset, lower3, ez_at_offset):
        arr, keep = _ptr_array(lower3)
        ez = np.ascontiguousarray(ez_at_offset, dtype=np.float64)
        self._check(self._fn("window_begin_centered")(self.h, step, offset, arr, _ptr(ez)))

    def window_set_inlet(self, offset, step1, ex, ey):
        ex = np.ascontiguousarray(ex, dtype=np.float64)
        ey = np.ascontiguousarray(ey, dtype=np.float64)
        self._check(self._fn("window_set_inlet_at")(self.h, offset, step1, _ptr(ex), _ptr(ey)))

    def window_shift(self, step, offset, planes6):
        arr, keep = _ptr_array(planes6)
        self._check(self._fn("window_shift")(self.h, step, offset, arr))

    def enable_ledger(self, enabled=True):
        """MovingWindowFields::enableEnergyLedger."""
        self._check(self._fn("window_enable_ledger")(self.h, int(enabled)))

    def finish_ledger_step(self):
        """MovingWindowFields::finishEnergyLedgerStep (after the commit)."""
        self._check(self._fn("window_finish_ledger_step")(self.h))

    def ledger(self):
        """(inlet overwrite, inserted, cropped, remap, outward x6, overwrite
        events, flux steps, remap steps) of the window's energy ledger."""
        out = np.zeros(13)
        self._check(self._fn("window_ledger")(self.h, _ptr(out)))
        return out

    def window_preview(self, enabled):
        self._check(self._fn("window_preview")(self.h, int(enabled)))

    def window_undo_preview(self):
        self._check(self._fn("window_undo_preview")(self.h))

    def window_commit(self):
        self._check(self._fn("window_commit")(self.h))

    def energy_density(self, i, j, k):
        return self._fn("fields_energy_density")(self.h, i, j, k)

    def poynting(self, i, j, k, axis):
        return self._fn("fields_poynting")(self.h, i, j, k, axis)


def cowan(prefix, dx, dy, dz):
    lib = load(prefix)
    out = np.zeros(12)
    code = getattr(lib, prefix + "_cowan")(dx, dy, dz, _ptr(out))
    if code:
        raise CpuError(code, getattr(lib, prefix + "_last_error")().decode())
    return out


def seeded_fields(fields, seed, components=range(6), scale=1.0):
    """Deterministic smooth+noisy fill (threaded_fields.cpp:29-41 style)."""
    rng = np.random.default_rng(seed)
    for a in components:
        v = fields.view(a)
        n = np.arange(v.size, dtype=np.float64).reshape(v.shape)
        v[...] = scale * (np.sin(0.245765 * n + 0.37 * a) + 0.25 * rng.standard_normal(v.shape))
        if a >= 3 and a < 6:
            v[...] /= C_SI


# ---------------------------------------------------------------------------
# Particle side.

_u64p = C.POINTER(C.c_uint64)


def _bind_particles(lib, prefix):
    P = prefix + "_"
    vp, sz = C.c_void_p, C.c_size_t
    f = lambda n: getattr(lib, P + n)  # noqa: E731
    f("coupling_new").argtypes = [vp, C.c_int, C.c_int, C.POINTER(vp)]
    f("coupling_free").argtypes = [vp]
    f("coupling_free").restype = None
    f("coupling_set_frame").argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double]
    f("coupling_add_magnet").argtypes = [vp, C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                         C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
    f("coupling_set_cpml_cells").argtypes = [vp, C.POINTER(sz)]
    f("coupling_set_particles").argtypes = [vp, sz, _dp, _dp, _dp, _dp, _dp, _u64p]
    f("coupling_count").argtypes = [vp]
    f("coupling_count").restype = sz
    f("coupling_get_particles").argtypes = [vp, _dp, _dp]
    f("coupling_step").argtypes = [vp, _dp, C.c_double, C.c_uint, C.c_double, C.POINTER(C.c_uint), _dp, _dp]
    f("coupling_begin_fields").argtypes = [vp, _dp]
    f("shape_weights").argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.POINTER(C.c_int),
                                   C.POINTER(C.c_longlong), _dp, sz, C.POINTER(sz), C.POINTER(C.c_uint)]
    f("momentum").argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_uint)]
    f("prescribed_box").argtypes = [vp, _dp, C.c_double, _dp, _dp]
    f("current_deposit").argtypes = [vp, C.c_int, C.c_int, _dp, _dp, _dp, C.c_double]
    f("binomial_xz_magnetic").argtypes = [vp, _dp, _dp, _dp]
    f("magnetic_sample_open").argtypes = [vp, _dp, _dp, _dp]
    f("coupling_get_ids").argtypes = [vp, _u64p]
    f("coupling_carrier_count").argtypes = [vp]
    f("coupling_carrier_count").restype = sz
    f("coupling_get_carriers").argtypes = [vp, _dp, _dp, _dp, _dp, _u64p]
    f("coupling_counters").argtypes = [vp, _u64p, _dp]
    f("coupling_finish_window").argtypes = [vp, _dp]
    f("coupling_set_held").argtypes = [vp, C.POINTER(C.c_int), _dp]
    f("coupling_get_frozen").argtypes = [vp, C.POINTER(C.c_int)]
    f("coupling_set_activation").argtypes = [vp, C.c_double]
    f("coupling_set_hold_transverse").argtypes = [vp, C.c_int]
    f("coupling_set_pending").argtypes = [vp, C.c_double, C.c_int64, C.c_double, C.c_double, C.c_uint64,
                                          C.c_uint64]
    f("coupling_admit").argtypes = [vp, C.c_uint64, C.POINTER(sz)]
    f("coupling_pending_state").argtypes = [vp, _u64p]
    f("coupling_pending_near").argtypes = [vp, C.c_uint64, C.c_uint64, sz, _dp, _u64p, C.POINTER(sz)]
    f("repeat_held_addition").argtypes = [C.c_double, C.c_double, C.c_uint64, _dp]
    if prefix == "orc":
        f("window_set_charge_match").argtypes = [vp, C.c_int]
        f("coupling_complete_inlet").argtypes = [vp, _dp]
        f("coupling_set_reference_current").argtypes = [vp, _dp, sz, sz, sz, C.c_double]
        f("reference_current_points").argtypes = [vp, sz, sz, C.POINTER(sz)]
        f("reference_current_points").restype = sz
    else:
        f("window_set_charge_match").argtypes = [vp, vp, _dp]
    f("window_begin_centered").argtypes = [vp, sz, C.c_ssize_t, C.POINTER(_dp), _dp]
    f("boundary_query").argtypes = [vp, _dp, C.c_int, _dp, _dp, C.c_double, _dp]


def load_particles(prefix):
    lib = load(prefix)
    if not getattr(lib, "_particles_bound", False):
        _bind_particles(lib, prefix)
        lib._particles_bound = True
    return lib


def _v(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def repeat_held_addition(prefix, initial, increment, count):
    """repeatHeldAddition of the oracle (literal additions) or the reference build."""
    out = C.c_double()
    _call(prefix, "repeat_held_addition", float(initial), float(increment), int(count), C.byref(out))
    return out.value


def _call(prefix, name, *args):
    lib = load_particles(prefix)
    code = getattr(lib, prefix + "_" + name)(*args)
    if code:
        raise CpuError(code, getattr(lib, prefix + "_last_error")().decode())


def shape_weights(prefix, shape, smooth_x, geometry7, origin, begin, end, capacity=20000):
    axis = np.zeros(capacity, np.int32)
    ijk = np.zeros(3 * capacity, np.int64)
    w = np.zeros(capacity)
    n = C.c_size_t()
    seg = C.c_uint()
    g, o, b, e = _v(geometry7), _v(origin), _v(begin), _v(end)
    _call(prefix, "shape_weights", shape, int(smooth_x), _ptr(g), _ptr(o), _ptr(b), _ptr(e),
          axis.ctypes.data_as(C.POINTER(C.c_int)), ijk.ctypes.data_as(C.POINTER(C.c_longlong)), _ptr(w),
          capacity, C.byref(n), C.byref(seg))
    k = n.value
    return axis[:k].copy(), ijk[:3 * k].reshape(k, 3).copy(), w[:k].copy(), seg.value


def momentum(prefix, u, e, b, q, m, h, g0=0.0):
    uu = _v(u).copy()
    it = C.c_uint()
    _call(prefix, "momentum", _ptr(uu), _ptr(_v(e)), _ptr(_v(b)), q, m, h, g0, C.byref(it))
    return uu, it.value


class CpuCoupling:
    """ConservativeSourceCoupling + particle loop in the oracle or reference."""

    SHAPES = {"quintic": 0, "notch": 1, "binomial3": 2}

    def __init__(self, fields: CpuFields, shape="binomial3", smooth_x=True, cpml_cells=(0, 0, 0)):
        self.prefix = fields.P[:-1]
        self.lib = load_particles(self.prefix)
        self.fields = fields
        h = C.c_void_p()
        _call(self.prefix, "coupling_new", fields.h, self.SHAPES[shape], int(smooth_x), C.byref(h))
        self.h = h
        cells = (C.c_size_t * 3)(*cpml_cells)
        _call(self.prefix, "coupling_set_cpml_cells", self.h, cells)

    def __del__(self):
        if getattr(self, "h", None):
            getattr(self.lib, self.prefix + "_coupling_free")(self.h)
            self.h = None

    def set_frame(self, gamma, lab_t0=0.0, lab_z0=0.0, box_z0=0.0):
        _call(self.prefix, "coupling_set_frame", self.h, gamma, lab_t0, lab_z0, box_z0)

    def add_magnet(self, kind, center, length, period=0.0, peak=0.0, angle=0.0, gaussian_fringe=False,
                   cutoff=1e-9, gradient=0.0, end_ramp=0.0, phase=0.0):
        t = {"undulator": 0, "dipole": 1, "quadrupole": 2}[kind]
        _call(self.prefix, "coupling_add_magnet", self.h, t, _ptr(_v(center)), length, period, peak, angle,
              int(gaussian_fringe), cutoff, gradient, end_ramp, phase)

    def set_particles(self, x, u, q, m, w, ids):
        x, u = _v(x), _v(u)
        q, m, w = _v(q), _v(m), _v(w)
        ids = np.ascontiguousarray(ids, dtype=np.uint64)
        _call(self.prefix, "coupling_set_particles", self.h, len(q), _ptr(x.ravel()), _ptr(u.ravel()), _ptr(q),
              _ptr(m), _ptr(w), ids.ctypes.data_as(_u64p))

    def particles(self):
        n = getattr(self.lib, self.prefix + "_coupling_count")(self.h)
        x = np.zeros((n, 3))
        u = np.zeros((n, 3))
        _call(self.prefix, "coupling_get_particles", self.h, _ptr(x), _ptr(u))
        return x, u

    def ids(self):
        n = getattr(self.lib, self.prefix + "_coupling_count")(self.h)
        ids = np.zeros(n, np.uint64)
        _call(self.prefix, "coupling_get_ids", self.h, ids.ctypes.data_as(_u64p))
        return ids

    def carriers(self):
        """(x, v, q, weight, id) of the CPML carriers in list order."""
        n = getattr(self.lib, self.prefix + "_coupling_carrier_count")(self.h)
        x, v = np.zeros((n, 3)), np.zeros((n, 3))
        q, w = np.zeros(n), np.zeros(n)
        ids = np.zeros(n, np.uint64)
        _call(self.prefix, "coupling_get_carriers", self.h, _ptr(x), _ptr(v), _ptr(q), _ptr(w),
              ids.ctypes.data_as(_u64p))
        return x, v, q, w, ids

    def counters(self):
        """Cumulative counts[20] and charges[20]: CPML entries per face (0-5),
        particle outer exits (6-11), carrier outer exits (12-17), window
        exits (18) and cropped carriers (19)."""
        counts = np.zeros(20, np.uint64)
        charges = np.zeros(20)
        _call(self.prefix, "coupling_counters", self.h, counts.ctypes.data_as(_u64p), _ptr(charges))
        return counts, charges

    def enable_charge_match(self, reference_origin):
        """Charge-matched inlet completion with a centered normal reference."""
        if self.prefix == "orc":
            _call(self.prefix, "window_set_charge_match", self.fields.h, 1)
        else:
            _call(self.prefix, "window_set_charge_match", self.fields.h, self.h, _ptr(_v(reference_origin)))

    def complete_inlet(self, new_origin):
        """The oracle completes the inlet after a shift; the reference build
        does it inside its shift."""
        if self.prefix == "orc":
            _call(self.prefix, "coupling_complete_inlet", self.h, _ptr(_v(new_origin)))

    def set_reference_current(self, table, guards, beta):
        t = _v(table)
        _call(self.prefix, "coupling_set_reference_current", self.h, _ptr(t.ravel()), t.shape[0], guards[0],
              guards[1], beta)

    def reference_points(self, guards):
        lib = load_particles(self.prefix)
        n = getattr(lib, self.prefix + "_reference_current_points")(self.h, guards[0], guards[1], None)
        ij = np.zeros(2 * n, dtype=np.uint64)
        getattr(lib, self.prefix + "_reference_current_points")(self.h, guards[0], guards[1],
                                                               ij.ctypes.data_as(C.POINTER(C.c_size_t)))
        return ij.reshape(n, 2).astype(np.int64)

    def set_held(self, frozen, beta):
        fr = np.ascontiguousarray(frozen, dtype=np.int32)
        _call(self.prefix, "coupling_set_held", self.h, fr.ctypes.data_as(C.POINTER(C.c_int)), _ptr(_v(beta)))

    def frozen(self):
        n = getattr(self.lib, self.prefix + "_coupling_count")(self.h)
        fr = np.zeros(n, dtype=np.int32)
        _call(self.prefix, "coupling_get_frozen", self.h, fr.ctypes.data_as(C.POINTER(C.c_int)))
        return fr

    def set_activation(self, lab_z):
        _call(self.prefix, "coupling_set_activation", self.h, lab_z)

    def set_hold_transverse(self, hold):
        _call(self.prefix, "coupling_set_hold_transverse", self.h, int(bool(hold)))

    def set_pending(self, reference_origin_z, initial_offset, dz, held_distance_per_beta, maximum_step,
                    initial_step=0):
        """PendingParticleQueue over the particles (after set_held)."""
        _call(self.prefix, "coupling_set_pending", self.h, reference_origin_z, initial_offset, dz,
              held_distance_per_beta, maximum_step, initial_step)

    def admit(self, step):
        n = C.c_size_t()
        _call(self.prefix, "coupling_admit", self.h, step, C.byref(n))
        return n.value

    def pending_state(self):
        """(waiting, next due step, admitted)."""
        out = np.zeros(3, np.uint64)
        _call(self.prefix, "coupling_pending_state", self.h, out.ctypes.data_as(_u64p))
        return tuple(int(v) for v in out)

    def pending_near(self, step, look_ahead):
        """(z, ids) of the waiting particles due by step + look_ahead."""
        n = C.c_size_t()
        _call(self.prefix, "coupling_pending_near", self.h, step, look_ahead, 0, None, None, C.byref(n))
        z = np.zeros(n.value)
        ids = np.zeros(n.value, np.uint64)
        if n.value:
            _call(self.prefix, "coupling_pending_near", self.h, step, look_ahead, n.value, _ptr(z),
                  ids.ctypes.data_as(_u64p), C.byref(n))
        return z, ids

    def finish_window(self, new_origin):
        _call(self.prefix, "coupling_finish_window", self.h, _ptr(_v(new_origin)))

    def boundary_query(self, origin, kind, a, b, duration=0.0):
        out = np.zeros(6)
        _call(self.prefix, "boundary_query", self.h, _ptr(_v(origin)), kind, _ptr(_v(a)), _ptr(_v(b)), duration,
              _ptr(out))
        return out

    def step(self, origin, time, substeps=1, exit_lab_z=float("inf")):
        it = C.c_uint()
        force = C.c_double()
        work = C.c_double()
        _call(self.prefix, "coupling_step", self.h, _ptr(_v(origin)), time, substeps, exit_lab_z, C.byref(it),
              C.byref(force), C.byref(work))
        return it.value, force.value, work.value

    def particle_energy(self, time):
        """Energy-ledger particle sums (simulation.cpp:3591-3625) at box time."""
        out = np.zeros(9)
        _call(self.prefix, "coupling_particle_energy", self.h, C.c_double(time), _ptr(out))
        return out

    def prescribed_box(self, pos, t):
        e = np.zeros(3)
        b = np.zeros(3)
        _call(self.prefix, "prescribed_box", self.h, _ptr(_v(pos)), t, _ptr(e), _ptr(b))
        return e, b


def current_deposit(fields: CpuFields, shape, smooth_x, origin, a, b, charge):
    prefix = fields.P[:-1]
    _call(prefix, "current_deposit", fields.h, CpuCoupling.SHAPES[shape], int(smooth_x), _ptr(_v(origin)),
          _ptr(_v(a)), _ptr(_v(b)), charge)


def binomial_xz_magnetic(fields: CpuFields, origin, point):
    out = np.zeros(3)
    _call(fields.P[:-1], "binomial_xz_magnetic", fields.h, _ptr(_v(origin)), _ptr(_v(point)), _ptr(out))
    return out


def magnetic_sample_open(fields: CpuFields, origin, point):
    out = np.zeros(3)
    _call(fields.P[:-1], "magnetic_sample_open", fields.h, _ptr(_v(origin)), _ptr(_v(point)), _ptr(out))
    return out


# ---------------------------------------------------------------------------
# Seeded 6D Gaussian bunch (particle_initializer.cpp:24-107: splitmix64 +
# Box-Muller per particle id), used for synthetic workloads.

_M64 = (1 << 64) - 1


def splitmix64(v):
    v = (v + 0x9E3779B97F4A7C15) & _M64
    v = ((v ^ (v >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    v = ((v ^ (v >> 27)) * 0x94D049BB133111EB) & _M64
    return v ^ (v >> 31)


# Shaped vertex charge of a source shape (the charge whose change the
# shape's current conserves): binomial3-x -- x nodeWeights x linear y x
# binomial3 z basis; quintic -- linear x, y x the compensated quintic z basis
# (longitudinal_shape.h:25-55). Positions relative to the grid origin.
def quintic_basis(z):
    cell = np.floor(z)
    f = z - cell
    n = [1.0, 0.0, 0.0, 0.0, 0.0, 0.0]
    for j in range(1, 6):
        saved = 0.0
        for k in range(j):
            t = n[k] / j
            n[k] = saved + (k + 1 - f) * t
            saved = (f + j - k - 1) * t
        n[j] = saved
    p = 5.0
    a = p / 6.0
    b = 3 * p / 40.0 + p * (p - 1) / 72.0
    d = 5 * p / 112.0 + p * (p - 1) / 80.0 + p * (p - 1) * (p - 2) / 1296.0
    c = [-d / 64.0, b / 16.0 + 3 * d / 32.0, -a / 4.0 - b / 4.0 - 15 * d / 64.0, 1 + a / 2.0 + 3 * b / 8.0 + 5 * d / 16.0,
         -a / 4.0 - b / 4.0 - 15 * d / 64.0, b / 16.0 + 3 * d / 32.0, -d / 64.0]
    v = np.zeros(12)
    for j in range(6):
        for k in range(7):
            v[j + k] += n[j] * c[k]
    return int(cell) - 5, v


def shaped_node_weight(shape, z, node):
    """longitudinal_shape::nodeWeight (longitudinal_shape.h:56) for the
    linear-binomial3 and compensated quintic shapes."""
    if shape == "binomial3":
        cz = np.floor(z)
        fz = z - cz
        k0 = int(cz) - 1
        w = [0.25 * (1 - fz), 0.5 * (1 - fz) + 0.25 * fz, 0.25 * (1 - fz) + 0.5 * fz, 0.25 * fz]
    else:
        k0, w = quintic_basis(z)
    return w[node - k0] if 0 <= node - k0 < len(w) else 0.0


def shaped_charge(shape, x, q, dims, h):
    nx, ny, nz = dims
    rho = np.zeros((nz + 1, ny + 1, nx + 1))
    inv = 1.0 / (h[0] * h[1] * h[2])
    for p in range(len(q)):
        lx, ly, lz = x[p, 0] / h[0], x[p, 1] / h[1], x[p, 2] / h[2]
        i, j = int(np.floor(lx)), int(np.floor(ly))
        fx, fy = lx - i, ly - j
        if shape == "binomial3":
            wx, i0 = [0.25 * (1 - fx), 0.5 - 0.25 * fx, 0.25 + 0.25 * fx, 0.25 * fx], i - 1
            cz = np.floor(lz)
            fz = lz - cz
            wz, k0 = [0.25 * (1 - fz), 0.5 * (1 - fz) + 0.25 * fz, 0.25 * (1 - fz) + 0.5 * fz, 0.25 * fz], int(cz) - 1
        else:
            wx, i0 = [1 - fx, fx], i
            k0, wz = quintic_basis(lz)
        for c, w3 in enumerate(wz):
            for b, w2 in enumerate([1 - fy, fy]):
                for a, w1 in enumerate(wx):
                    kk, jj, ii = k0 + c, j + b, i0 + a
                    if 0 <= kk <= nz and 0 <= jj <= ny and 0 <= ii <= nx:
                        rho[kk, jj, ii] += q[p] * inv * w1 * w2 * w3
    return rho


def continuity(rho0, rho1, jx, jy, jz, dt, h):
    rate = (rho1 - rho0) / dt
    div = np.zeros_like(rho0)
    div[:, :, :-1] += jx / h[0]
    div[:, :, 1:] -= jx / h[0]
    div[:, :-1, :] += jy / h[1]
    div[:, 1:, :] -= jy / h[1]
    div[:-1, :, :] += jz / h[2]
    div[1:, :, :] -= jz / h[2]
    return np.abs(rate + div).max(), np.abs(rate).max()
