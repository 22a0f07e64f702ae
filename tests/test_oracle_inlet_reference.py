"""The oracle's inlet reference (oracle/orc_inlet_reference.hpp), a
restatement of BlockedCowanReference (blocked_cowan_reference.cpp) and
movingWindowFieldValue (simulation_moving_window.cpp:280-294). The
reference translation unit needs FFTW3, which this image lacks, so the
restatement is pinned without a reference build:

* against the Cowan initializer: on the same grid, CIC source, beta and
  static guards the reference's fields are those of the production Cowan
  initializer (blocked_cowan_reference.h:68-70); the oracle's initializer
  (bitwise equal to the reference build's EBGaussFieldInitializer) solved by
  conjugate gradients to 1e-13 gives E and B within 1e-11 of the modal
  solve's;
* the potential satisfies the 27-point Cowan operator with the translation
  term (assembled here from the Cowan coefficients of the reference build)
  to round-off on every interior vertex, and vanishes on the Dirichlet
  planes;
* the shaped source (linear-binomial3 with the x-binomial3 transverse
  weights, the quintic basis) conserves the total charge and moves its
  centroid with the particles;
* movingWindowFieldValue: exact at whole-plane shifts, linear between,
  the reference's out-of-range error beyond the prepared planes;
* a planted sign error in the translation term is caught by the operator
  check.
"""
import ctypes as C

import numpy as np
import pytest

import harness as H
from test_oracle_initial_field import bunch_case, gauss

_dp = C.POINTER(C.c_double)
_sz = C.c_size_t
SHAPES = {"cic": -1, "quintic": 0, "binomial3": 2}


class OrcReference:
    def __init__(self, dims, spacing, dt, origin, x, q, beta, guards, block_layers=16, shape="cic",
                 x_binomial=False):
        lib = H.load("orc")
        lib.orc_inlet_reference_new.argtypes = [C.POINTER(_sz), _dp, _dp, _sz, _dp, _dp, C.c_double, C.POINTER(_sz),
                                                _sz, C.c_int, C.c_int, C.POINTER(C.c_void_p), _dp]
        lib.orc_inlet_reference_free.argtypes = [C.c_void_p]
        lib.orc_inlet_reference_free.restype = None
        lib.orc_inlet_reference_array.argtypes = [C.c_void_p, C.c_int, _dp, _sz]
        lib.orc_inlet_reference_moving.argtypes = [C.c_void_p, C.c_uint, C.c_ssize_t, _sz, C.c_double, _dp, _sz]
        self.lib = lib
        self.dims = dims
        cells = (_sz * 3)(*dims)
        sp = np.array([*spacing, dt])
        o = np.ascontiguousarray(origin, dtype=np.float64)
        xx = np.ascontiguousarray(x, dtype=np.float64)
        qq = np.ascontiguousarray(q, dtype=np.float64)
        g = (_sz * 3)(*guards)
        self.report = np.zeros(4)
        h = C.c_void_p()
        rc = lib.orc_inlet_reference_new(cells, H._ptr(sp), H._ptr(o), len(qq), H._ptr(xx), H._ptr(qq), beta, g,
                                         block_layers, SHAPES[shape], int(x_binomial), C.byref(h),
                                         H._ptr(self.report))
        if rc:
            raise H.CpuError(rc, lib.orc_last_error().decode())
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_inlet_reference_free(self.h)

    def array(self, which, shape):
        out = np.zeros(int(np.prod(shape)))
        rc = self.lib.orc_inlet_reference_array(self.h, which, H._ptr(out), out.size)
        if rc:
            raise H.CpuError(rc, self.lib.orc_last_error().decode())
        return out.reshape(shape)

    def component(self, a):
        return self.array(a, H.component_shape(*self.dims, a))

    def phi(self):
        nx, ny, nz = self.dims
        return self.array(6, (nz + 1, ny + 1, nx + 1))

    def moving(self, a, k, step, translation_beta):
        shape = H.component_shape(*self.dims, a)[1:]
        out = np.zeros(int(np.prod(shape)))
        rc = self.lib.orc_inlet_reference_moving(self.h, a, k, step, translation_beta, H._ptr(out), out.size)
        if rc:
            raise H.CpuError(rc, self.lib.orc_last_error().decode())
        return out.reshape(shape)


DIMS, SPACING = (16, 16, 32), (2.5e-6, 2.5e-6, 1e-6)
DT = SPACING[2] / H.C_SI
GUARDS = (3, 3, 2)


def case(n=300, uz=0.3):
    return bunch_case(DIMS, SPACING, n, uz)


def test_fields_equal_the_cowan_initializer():
    x, u, q, m = case()
    orc = H.CpuFields("orc", *DIMS, *SPACING, DT, 1)
    rep = gauss("orc", orc, np.zeros(3), x, u, q, m, "cowan-relativistic-poisson", GUARDS, tol=1e-13, max_it=20000)
    beta = rep[4]
    assert list(rep[14:17]) == list(GUARDS) and rep[17] == 0  # no transverse padding
    ref = OrcReference(DIMS, SPACING, DT, np.zeros(3), x, q, beta, GUARDS)
    for a in range(6):
        r = ref.component(a)
        o = orc.view(a)
        scale = np.abs(o).max()
        err = np.abs(r - o).max() / scale if scale > 0 else np.abs(r).max()
        print(a, f"{err:.2e}", scale)
        assert err <= 1e-11, (a, err)
    # the source charge is the particles' and the modes are positive definite
    assert ref.report[2] == pytest.approx(q.sum(), rel=1e-14)
    assert ref.report[3] == pytest.approx(q.sum(), rel=1e-12)
    assert ref.report[0] > 0 and ref.report[1] > 0


def cowan_apply(phi, dims, spacing, beta, cw, translation_sign=1.0):
    """27-point Cowan operator with the translation term on interior
    vertices: sum over axes of the smoothed second difference, assembled
    from the reference build's Cowan coefficients."""
    alpha, bcoef = cw[0:3], cw[3:6]
    delta_xy, delta_yz, delta_zx = cw[6], cw[7], cw[8]
    deltas = [delta_yz, delta_zx, delta_xy]
    out = np.zeros_like(phi)
    k, j, i = np.meshgrid(np.arange(1, dims[2]), np.arange(1, dims[1]), np.arange(1, dims[0]), indexing="ij")
    p = phi
    for axis in range(3):
        ta, tb = (axis + 1) % 3, (axis + 2) % 3
        inv = 1.0 / spacing[axis] ** 2
        for a in (-1, 0, 1):
            for b in (-1, 0, 1):
                w = (alpha[axis] if b == 0 else bcoef[tb]) if a == 0 else (bcoef[ta] if b == 0 else deltas[axis])
                d = [0, 0, 0]
                d[ta] += a
                d[tb] += b
                c0 = p[k + d[2], j + d[1], i + d[0]]
                dm = list(d)
                dm[axis] -= 1
                dp = list(d)
                dp[axis] += 1
                cm = p[k + dm[2], j + dm[1], i + dm[0]]
                cp = p[k + dp[2], j + dp[1], i + dp[0]]
                out[k, j, i] += w * inv * (2 * c0 - cm - cp)
    t = translation_sign * (beta * H.C_SI) ** 2 * 8.8541878128e-12 * 1.25663706212e-6 / spacing[2] ** 2
    out[k, j, i] -= t * (2 * p[k, j, i] - p[k - 1, j, i] - p[k + 1, j, i])
    return out


def cowan_coefficients():
    c = H.cowan("orc", *SPACING)
    return c


WIDE = (24, 24, 48)  # room for the wide shapes' support inside the guards


@pytest.mark.parametrize("shape,x_binomial", [("cic", False), ("binomial3", True), ("binomial3", False),
                                              ("quintic", False)])
def test_potential_solves_the_cowan_operator(shape, x_binomial):
    dims = WIDE
    x, u, q, m = bunch_case(dims, SPACING, 300, 0.3)
    beta = 0.29
    ref = OrcReference(dims, SPACING, DT, np.zeros(3), x, q, beta, GUARDS, shape=shape, x_binomial=x_binomial)
    phi = ref.phi()
    rho = ref.array(7, (dims[2] - 2 * GUARDS[2] - 1, dims[1] - 2 * GUARDS[1] - 1, dims[0] - 2 * GUARDS[0] - 1))
    cw = cowan_coefficients()
    image = cowan_apply(phi, dims, SPACING, beta, cw)
    lo = GUARDS
    interior = image[lo[2] + 1:dims[2] - lo[2], lo[1] + 1:dims[1] - lo[1], lo[0] + 1:dims[0] - lo[0]]
    target = rho / 8.8541878128e-12
    err = np.abs(interior - target).max() / np.abs(target).max()
    print(shape, "operator residual", err)
    assert err <= 1e-12
    # zero on and outside the Dirichlet planes
    mask = np.ones_like(phi, dtype=bool)
    mask[lo[2] + 1:dims[2] - lo[2], lo[1] + 1:dims[1] - lo[1], lo[0] + 1:dims[0] - lo[0]] = False
    assert not np.any(phi[mask])
    # the source keeps the particles' charge and centroid
    h = np.prod(SPACING)
    assert rho.sum() * h == pytest.approx(q.sum(), rel=1e-12)
    zc = (np.arange(rho.shape[0]) + lo[2] + 1) * SPACING[2]
    assert np.sum(rho.sum(axis=(1, 2)) * zc) / rho.sum() == pytest.approx(np.sum(q * x[:, 2]) / q.sum(), rel=1e-9)


def test_planted_translation_sign_is_caught():
    x, u, q, m = case()
    beta = 0.29
    ref = OrcReference(DIMS, SPACING, DT, np.zeros(3), x, q, beta, GUARDS)
    phi = ref.phi()
    rho = ref.array(7, (DIMS[2] - 2 * GUARDS[2] - 1, DIMS[1] - 2 * GUARDS[1] - 1, DIMS[0] - 2 * GUARDS[0] - 1))
    cw = cowan_coefficients()
    # the operator with the translation term's sign flipped does not fit
    image = cowan_apply(phi, DIMS, SPACING, beta, cw, translation_sign=-1.0)
    lo = GUARDS
    interior = image[lo[2] + 1:DIMS[2] - lo[2], lo[1] + 1:DIMS[1] - lo[1], lo[0] + 1:DIMS[0] - lo[0]]
    target = rho / 8.8541878128e-12
    assert np.abs(interior - target).max() / np.abs(target).max() > 1e-3


def test_moving_window_value():
    x, u, q, m = case()
    ref = OrcReference(DIMS, SPACING, DT, np.zeros(3), x, q, 0.29, GUARDS)
    for a in range(6):
        comp = ref.component(a)
        nz = comp.shape[0]
        # a whole-plane shift reads the plane itself
        assert np.array_equal(ref.moving(a, 10, 4, 0.5), comp[8])
        # between planes: linear
        v = ref.moving(a, 10, 3, 0.5)
        assert np.allclose(v, 0.5 * comp[8] + 0.5 * comp[9], rtol=1e-15, atol=1e-300)
        # the last plane itself, then beyond it
        assert np.array_equal(ref.moving(a, nz - 1, 0, 0.3), comp[nz - 1])
        with pytest.raises(H.CpuError, match="Moving reference query outside prepared native field"):
            ref.moving(a, nz, 0, 0.3)
        with pytest.raises(H.CpuError, match="outside prepared native field"):
            ref.moving(a, 1, 5, 0.5)
