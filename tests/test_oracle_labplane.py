"""The oracle's laboratory field plane (oracle/orc_detectors.hpp), pinned
without the reference's detector translation unit (its sampler lives in an
anonymous namespace of lab_detectors.cpp, which also needs HDF5):

* the BoostFrameTransform members it calls (boxZFromLabZAndBoxT,
  labTFromLabZBoxT, fieldsBoxToLab; boostframe.cpp:199-285) bitwise
  against the reference build;
* the clamped trilinear stencil by closed forms: a field linear in the
  position of each component's own Yee nodes is sampled exactly at the
  cell centres of the plane, for any plane inside [0.5, nz - 0.5] cells;
  below and above, the half-node components hold their end plane
  (interpolationAxis clamps, lab_detectors.cpp:39-58);
* the E/B mix of the lab transform against its closed form, and a planted
  half-cell slip of the stencil caught by the linear field;
* the cadence of sampleFieldPlanes (lab_detectors.cpp:2439-2484): with a
  rhythm of three steps' lab time (less one part in 10^9) every third
  step samples; with a rhythm far below one step's every step does; a
  plane off the grid never samples; a rhythm that is not positive is
  rejected as the reference's configuration rejects it (config.cpp:
  1713-1717).
"""
import ctypes as C

import numpy as np
import pytest

import harness as H

_dp = C.POINTER(C.c_double)


def _boost(prefix):
    lib = H.load(prefix)
    fn = getattr(lib, prefix + "_boost_lab")
    fn.argtypes = [C.c_double] * 6 + [_dp, _dp, _dp]
    return fn


def boost_lab(prefix, frame, z_lab, t_box, e, b):
    out = np.zeros(8)
    e = np.ascontiguousarray(e, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    assert _boost(prefix)(*frame, z_lab, t_box, H._ptr(e), H._ptr(b), H._ptr(out)) == 0
    return out


class OrcPlane:
    def __init__(self, z_lab, rhythm):
        lib = H.load("orc")
        self.h = None
        lib.orc_lab_plane_new.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_void_p)]
        lib.orc_lab_plane_free.argtypes = [C.c_void_p]
        lib.orc_lab_plane_free.restype = None
        lib.orc_lab_plane_sample.argtypes = [C.c_void_p, C.c_void_p, _dp, C.c_double, _dp, _dp, _dp,
                                             C.POINTER(C.c_int)]
        self.lib = lib
        h = C.c_void_p()
        assert lib.orc_lab_plane_new(z_lab, rhythm, C.byref(h)) == 0
        self.h = h

    def __del__(self):
        if self.h:
            self.lib.orc_lab_plane_free(self.h)

    def sample(self, fields, frame, t_box, origin):
        rec = np.zeros((fields.ny, fields.nx, 6))
        t_lab = C.c_double(0.0)
        taken = C.c_int(0)
        f = np.ascontiguousarray(frame, dtype=np.float64)
        o = np.ascontiguousarray(origin, dtype=np.float64)
        assert self.lib.orc_lab_plane_sample(self.h, fields.h, H._ptr(f), t_box, H._ptr(o), H._ptr(rec.reshape(-1)),
                                             C.byref(t_lab), C.byref(taken)) == 0
        return bool(taken.value), t_lab.value, rec


@pytest.mark.skipif(not H.have_ref(), reason="reference build unavailable")
def test_boost_members_bitwise_against_reference_build():
    rng = np.random.default_rng(11)
    for _ in range(200):
        frame = (rng.uniform(1.0, 500.0), rng.normal() * 1e-9, rng.normal() * 0.1, rng.normal() * 1e-3)
        z, t = rng.normal() * 0.5, rng.uniform(0, 1e-9)
        e, b = rng.normal(size=3) * 1e6, rng.normal(size=3)
        assert np.array_equal(boost_lab("orc", frame, z, t, e, b), boost_lab("ref", frame, z, t, e, b))


def linear_grid(dims, spacing, coef):
    """Oracle grid whose component a holds coef[a] . (1, X, Y, Z) at its own
    Yee node positions (E: edge centres, B: face centres)."""
    nx, ny, nz = dims
    f = H.CpuFields("orc", nx, ny, nz, *spacing, spacing[2] / H.C_SI, 1)
    for a in range(6):
        shape = H.component_shape(nx, ny, nz, a)
        k, j, i = np.meshgrid(*[np.arange(n, dtype=float) for n in shape], indexing="ij")
        axis = a % 3
        half = [0.0, 0.0, 0.0]
        for d in range(3):
            half[d] = 0.5 if (d == axis) == (a < 3) else 0.0
        X, Y, Z = (i + half[0]) * spacing[0], (j + half[1]) * spacing[1], (k + half[2]) * spacing[2]
        f.view(a)[...] = coef[a][0] + coef[a][1] * X + coef[a][2] * Y + coef[a][3] * Z
    return f


def lab_of(frame, eb):
    """fieldsBoxToLab closed form (boostframe.cpp:267-285)."""
    g = frame[0]
    gb = np.sqrt(g - 1.0) * np.sqrt(g + 1.0)
    e, b = eb[..., :3], eb[..., 3:]
    out = np.empty_like(eb)
    out[..., 0] = g * e[..., 0] + gb * H.C_SI * b[..., 1]
    out[..., 1] = g * e[..., 1] - gb * H.C_SI * b[..., 0]
    out[..., 2] = e[..., 2]
    out[..., 3] = g * b[..., 0] - gb / H.C_SI * e[..., 1]
    out[..., 4] = g * b[..., 1] + gb / H.C_SI * e[..., 0]
    out[..., 5] = b[..., 2]
    return out


def plane_time_for(frame, z_lab, box_z):
    """Box time at which the lab plane z_lab sits at box z (inverse of
    boxZFromLabZAndBoxT)."""
    g, _, lz0, bz0 = frame
    beta = np.sqrt(g - 1.0) * np.sqrt(g + 1.0) / g
    return (bz0 + (z_lab - lz0) / g - box_z) / (beta * H.C_SI)


def test_linear_fields_sampled_exactly_at_cell_centres():
    dims, spacing = (9, 7, 20), (3e-6, 2e-6, 1e-6)
    rng = np.random.default_rng(3)
    coef = rng.normal(size=(6, 4)) * np.array([1.0, 1e5, 1e5, 1e5])
    f = linear_grid(dims, spacing, coef)
    frame = (7.0, 0.0, 0.0, 0.0)
    origin = np.array([0.0, 0.0, -4e-6])
    z_lab = 1e-3
    for cz in (0.5, 0.75, 3.2, 10.0, 19.4, 19.5):  # plane z in cells above the origin
        box_z = origin[2] + cz * spacing[2]
        t = plane_time_for(frame, z_lab, box_z)
        p = OrcPlane(z_lab, 1e-30)
        taken, t_lab, rec = p.sample(f, frame, t, origin)
        assert taken
        Xc = (np.arange(dims[0]) + 0.5) * spacing[0]
        Yc = (np.arange(dims[1]) + 0.5) * spacing[1]
        X, Y = np.meshgrid(Xc, Yc)
        Z = cz * spacing[2]
        box = np.stack([coef[a][0] + coef[a][1] * X + coef[a][2] * Y + coef[a][3] * Z for a in range(6)], axis=-1)
        want = lab_of(frame, box)
        scale = np.abs(want).max(axis=(0, 1))
        assert np.all(np.abs(rec - want) <= 1e-12 * scale), (cz, np.abs(rec - want).max(axis=(0, 1)) / scale)
        g = frame[0]
        beta = np.sqrt(g - 1.0) * np.sqrt(g + 1.0) / g
        assert t_lab == pytest.approx(t / g + beta * z_lab / H.C_SI, rel=1e-14)


def test_clamped_ends_and_a_planted_half_cell_slip():
    dims, spacing = (6, 5, 12), (2e-6, 2e-6, 1e-6)
    rng = np.random.default_rng(5)
    coef = rng.normal(size=(6, 4)) * np.array([1.0, 1e5, 1e5, 1e5])
    f = linear_grid(dims, spacing, coef)
    frame = (3.0, 0.0, 0.0, 0.0)
    origin = np.zeros(3)
    z_lab = 5e-4
    # a plane 0.2 cells above the origin: vertex components (Ex, Ey, Bz)
    # interpolate, the half-node ones (Ez, Bx, By) hold plane k = 0
    t = plane_time_for(frame, z_lab, 0.2 * spacing[2])
    taken, _, rec = OrcPlane(z_lab, 1e-30).sample(f, frame, t, origin)
    assert taken
    X, Y = np.meshgrid((np.arange(dims[0]) + 0.5) * spacing[0], (np.arange(dims[1]) + 0.5) * spacing[1])
    zs = [0.2, 0.2, 0.5, 0.5, 0.5, 0.2]
    box = np.stack([coef[a][0] + coef[a][1] * X + coef[a][2] * Y + coef[a][3] * zs[a] * spacing[2]
                    for a in range(6)], axis=-1)
    want = lab_of(frame, box)
    assert np.allclose(rec, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    # the same check against a stencil slipped by half a cell in x fails
    slipped = np.stack([coef[a][0] + coef[a][1] * (X + 0.5 * spacing[0]) + coef[a][2] * Y +
                        coef[a][3] * zs[a] * spacing[2] for a in range(6)], axis=-1)
    assert not np.allclose(rec, lab_of(frame, slipped), rtol=1e-9, atol=1e-9 * np.abs(want).max())


def test_cadence_every_third_step_and_off_grid_planes():
    dims, spacing = (4, 4, 40), (2e-6, 2e-6, 1e-6)
    f = linear_grid(dims, spacing, np.ones((6, 4)))
    frame = (5.0, 0.0, 0.0, 0.0)
    dt = spacing[2] / H.C_SI
    z_lab = 1e-5
    t0 = plane_time_for(frame, z_lab, 20 * spacing[2])  # the plane crosses the window centre
    step_lab = dt / frame[0]
    for rhythm, every in ((3 * step_lab * (1 - 1e-9), 3), (1e-3 * step_lab, 1)):
        p = OrcPlane(z_lab, rhythm)
        taken = [p.sample(f, frame, t0 + n * dt, np.zeros(3))[0] for n in range(12)]
        assert taken == [n % every == 0 for n in range(12)], (rhythm, taken)
    # a plane far beyond the grid never samples
    p = OrcPlane(1.0, step_lab)
    assert not any(p.sample(f, frame, t0 + n * dt, np.zeros(3))[0] for n in range(5))
    lib = H.load("orc")
    h = C.c_void_p()
    for bad in (0.0, -1.0, float("nan")):
        assert lib.orc_lab_plane_new(z_lab, bad, C.byref(h)) == 1
        assert b"rhythm must be positive" in lib.orc_last_error()
