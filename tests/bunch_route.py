"""The laboratory-plane input route in the oracle (prefix "orc") and the
reference build (prefix "ref"): bindings of the *_bunch_* entry points and
the frozen lab-plane injection sequence of Simulation's initialization
(simulation.cpp:1983-2245) over them, plus the plane-arrival events of a
seeded Gaussian bunch."""
import ctypes as C

import numpy as np

import harness as H
from aprl_amd import bunch

_dp = C.POINTER(C.c_double)
_sz = C.c_size_t


def _lib(prefix):
    lib = H.load(prefix)
    if getattr(lib, "_bunch_bound", False):
        return lib
    P = prefix + "_bunch_"
    getattr(lib, P + "project").argtypes = [_sz, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_int, _dp]
    getattr(lib, P + "place").argtypes = [_sz, _dp, C.c_int, C.c_double, C.c_double, C.c_double, _dp]
    getattr(lib, P + "frozen_origin").argtypes = [_sz, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                                  C.c_int, _dp]
    getattr(lib, P + "boost").argtypes = [_sz, _dp, _dp, _dp, C.c_double, C.c_int, C.c_double, _dp]
    getattr(lib, P + "freeze").argtypes = [_sz, _dp, _dp, _dp, C.c_double, C.c_int, _dp]
    lib._bunch_bound = True
    return lib


def _run(prefix, name, *args):
    lib = _lib(prefix)
    code = getattr(lib, prefix + "_bunch_" + name)(*args)
    if code:
        raise H.CpuError(code, getattr(lib, prefix + "_last_error")().decode())


def _p(a):
    return H._ptr(a)


def project(prefix, x, u, w, input_plane, reference, offset=0.0, hold=True):
    """In place on x (n, 3); returns the 8-entry report."""
    r = np.zeros(8)
    _run(prefix, "project", len(x), _p(x), _p(u), _p(w), input_plane, reference, offset, int(hold), _p(r))
    return r


def place(prefix, x, entrance, reference, margin):
    r = np.zeros(7)
    has = entrance is not None
    _run(prefix, "place", len(x), _p(x), int(has), float(entrance) if has else float("nan"), reference, margin,
         _p(r))
    return r


def frozen_origin(prefix, x, u, gamma_boost, input_plane, activation_plane, dz, hold=True):
    out = np.zeros(2)
    _run(prefix, "frozen_origin", len(x), _p(x), _p(u), gamma_boost, input_plane, activation_plane, dz, int(hold),
         _p(out))
    return out


def boost(prefix, x, u, frame4, snapshot_time=0.0, hold=True, input_plane=0.0):
    r = np.zeros(7)
    f = np.ascontiguousarray(frame4, dtype=np.float64)
    _run(prefix, "boost", len(x), _p(x), _p(u), _p(f), snapshot_time, int(hold), input_plane, _p(r))
    return r


def freeze(prefix, x, u, frame4, activation_plane, hold=True):
    beta = np.zeros(len(x))
    f = np.ascontiguousarray(frame4, dtype=np.float64)
    _run(prefix, "freeze", len(x), _p(x), _p(u), _p(f), activation_plane, int(hold), _p(beta))
    return beta


def plane_events(n, seed, sigma_xy, sigma_t, gamma_lab, sigma_gamma, emittance, electrons, center_xy=(0.0, 0.0)):
    """Plane-arrival events of a seeded 6D Gaussian: (x, y, arrival time
    offset) per row in ``x``, the lab proper velocity in ``u``, charges,
    masses, unit weights and ids 1..n (the same splitmix64 draws as the
    generated-Gaussian beam, axis 2 in seconds)."""
    u_lab = np.sqrt(gamma_lab ** 2 - 1.0)
    sig_u = [emittance / sigma_xy[0], emittance / sigma_xy[1], sigma_gamma * gamma_lab]
    x, u, q, m, ids = bunch.seeded_gaussian(n, seed, [center_xy[0], center_xy[1], 0.0],
                                            [sigma_xy[0], sigma_xy[1], sigma_t], [0.0, 0.0, u_lab], sig_u, electrons)
    return x, u, q, m, np.ones(n), ids


def frozen_injection(prefix, x, u, w, *, gamma_boost, input_plane, activation_plane, reference, dz, entrance,
                     offset=0.0):
    """Simulation's frozen lab-plane injection with the transverse hold
    (the only form a moving window accepts, config.cpp:2071-2076) on copies
    of the events: projection, placement, the frame origin at the earliest
    crossing, the boost to box time zero and the frozen flags. Returns
    (x_box, u_box, beta_held, frame4, reports)."""
    x = np.array(x, dtype=np.float64)
    u = np.array(u, dtype=np.float64)
    rp = project(prefix, x, u, np.asarray(w, dtype=np.float64), input_plane, reference, offset, True)
    rl = place(prefix, x, entrance, reference, gamma_boost * dz)
    origin = frozen_origin(prefix, x, u, gamma_boost, input_plane, activation_plane, dz, True)
    frame4 = np.array([gamma_boost, origin[0], origin[1], 0.0])
    rb = boost(prefix, x, u, frame4, 0.0, True, input_plane)
    beta = freeze(prefix, x, u, frame4, activation_plane, True)
    return x, u, beta, frame4, (rp, rl, origin, rb)
