"""Host-side mirror of the reference's field interfaces over the C-ABI.

``EBFieldGrid`` keeps the names of eb_field.h:201-357 (``ex()`` .. ``jz()``,
``advanceMagnetic``, ``advanceElectric``, ``clearCurrent``, ``setBoundary``)
and ``MovingWindowFields`` those of moving_window_fields.h:72-134. Component
accessors return host copies of the device arrays shaped (nz, ny, nx) in the
reference's x-fastest layout; ``set_component`` uploads.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _capi
from ._capi import check, dptr, lib, plane_array

EX, EY, EZ, BX, BY, BZ, JX, JY, JZ = range(9)
C_SI = 299792458.0
EPSILON0 = 8.8541878128e-12
MU0 = 1.25663706212e-6


@dataclass
class EBGridGeometry:
    """EBGridGeometry, eb_field.h:54-70. solver: 'yee' or 'cowan-z'."""
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    dt: float
    solver: str = "cowan-z"

    def c_struct(self) -> _capi.GridGeometry:
        return _capi.GridGeometry(self.nx, self.ny, self.nz, self.dx, self.dy, self.dz, self.dt,
                                  1 if self.solver == "cowan-z" else 0)


@dataclass
class EBCPMLParameters:
    """EBCPMLParameters, eb_cpml.h:12-24 (defaults eb_cpml.cpp:251-256)."""
    cells: tuple = (0, 0, 0)
    polynomialOrder: float = 3.0
    targetReflection: float = 1.0e-8
    kappaMax: float = 8.0
    alphaFraction: float = 0.0

    def c_struct(self) -> _capi.CpmlParameters:
        p = _capi.CpmlParameters()
        for a in range(3):
            p.cells[a] = int(self.cells[a])
        p.polynomial_order = self.polynomialOrder
        p.target_reflection = self.targetReflection
        p.kappa_max = self.kappaMax
        p.alpha_fraction = self.alphaFraction
        return p


def component_shape(g: EBGridGeometry, a: int) -> tuple:
    electric = a < 3 or a >= 6
    axis = a % 3
    cells = [g.nx, g.ny, g.nz]
    s = [cells[d] + ((d != axis) if electric else (d == axis)) for d in range(3)]
    return (s[2], s[1], s[0])


PRECISIONS = {"fp64": 0, "fp32-split": 1}


class EBFieldGrid:
    """Device-resident Yee grid (EBFieldGrid, eb_field.h:201-357).

    precision "fp64" is the plain form; "fp32-split" keeps fields, currents
    and CPML histories in FP32 (host arrays stay float64 and are rounded on
    upload)."""

    def __init__(self, geometry: EBGridGeometry, precision: str = "fp64"):
        self.geometry = geometry
        self.precision = precision
        self._g = geometry.c_struct()
        h = C.c_void_p()
        check(lib().aprl_fields_create(C.byref(self._g), PRECISIONS[precision], C.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().aprl_fields_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- components -------------------------------------------------------
    def component(self, a: int) -> np.ndarray:
        out = np.empty(component_shape(self.geometry, a), dtype=np.float64)
        check(lib().aprl_fields_download(self.handle, a, dptr(out), out.size))
        return out

    def component_plane(self, a: int, k: int) -> np.ndarray:
        """z plane k of component a, shape (ny_a, nx_a)."""
        shape = component_shape(self.geometry, a)[1:]
        out = np.empty(shape, dtype=np.float64)
        check(lib().aprl_fields_download_plane(self.handle, a, k, dptr(out), out.size))
        return out

    def set_component(self, a: int, values) -> None:
        arr = np.ascontiguousarray(values, dtype=np.float64)
        if arr.size != int(np.prod(component_shape(self.geometry, a))):
            raise ValueError("component size mismatch")
        check(lib().aprl_fields_upload(self.handle, a, dptr(arr.ravel()), arr.size))

    def ex(self): return self.component(EX)
    def ey(self): return self.component(EY)
    def ez(self): return self.component(EZ)
    def bx(self): return self.component(BX)
    def by(self): return self.component(BY)
    def bz(self): return self.component(BZ)
    def jx(self): return self.component(JX)
    def jy(self): return self.component(JY)
    def jz(self): return self.component(JZ)

    # -- boundary ---------------------------------------------------------
    def setBoundary(self, cpml: EBCPMLParameters, globalNz=None, localZCellOffset=0,
                    lowerZPhysical=True, upperZPhysical=True) -> None:
        """Install EBConvolutionalPML (eb_cpml.h:33-39) as the boundary."""
        p = cpml.c_struct()
        check(lib().aprl_fields_set_cpml(self.handle, C.byref(p), globalNz or self.geometry.nz,
                                         localZCellOffset, int(lowerZPhysical), int(upperZPhysical)))

    def checkpointHistory(self, index: int) -> np.ndarray:
        n = lib().aprl_fields_history_size(self.handle, index)
        out = np.empty(n, dtype=np.float64)
        check(lib().aprl_fields_history_download(self.handle, index, dptr(out) if n else None, n))
        return out

    def restoreHistory(self, index: int, values) -> None:
        arr = np.ascontiguousarray(values, dtype=np.float64)
        check(lib().aprl_fields_history_upload(self.handle, index, dptr(arr) if arr.size else None, arr.size))

    # -- stepping ---------------------------------------------------------
    def clearCurrent(self) -> None:
        check(lib().aprl_fields_clear_current(self.handle))

    def advanceMagnetic(self, halo=None) -> None:
        """halo: None, or six planes (lower Ex,Ey,Ez, upper Ex,Ey,Ez; None = PEC parity)."""
        arr, keep = plane_array(halo)
        check(lib().aprl_fields_advance_magnetic(self.handle, arr))

    def advanceElectric(self) -> None:
        check(lib().aprl_fields_advance_electric(self.handle))

    def advance(self) -> None:
        self.advanceMagnetic()
        self.advanceElectric()

    def synchronize(self) -> None:
        check(lib().aprl_fields_synchronize(self.handle))

    # -- timing -----------------------------------------------------------
    def interiorFieldEnergy(self, lo, hi):
        """Sum of fieldEnergyDensityCell * dV over the cells [lo, hi)
        (Simulation::localInteriorFieldEnergy, simulation.cpp:3308-3342)."""
        out = C.c_double()
        check(lib().aprl_fields_interior_energy(self.handle, _capi.size3(lo), _capi.size3(hi), C.byref(out)))
        return out.value

    def facePower(self, lo, hi):
        """Outward Poynting power through the six faces of the cell box
        [lo, hi) (Simulation::localInteriorBoundaryPower,
        simulation.cpp:3344-3432): [-x, +x, -y, +y, -z, +z]."""
        out = np.zeros(6)
        check(lib().aprl_fields_face_power(self.handle, _capi.size3(lo), _capi.size3(hi), dptr(out)))
        return out

    def profile(self, enable: bool) -> None:
        check(lib().aprl_profile_enable(self.handle, int(enable)))

    def profile_query(self, kernel: str):
        n = C.c_uint64()
        ms = C.c_double()
        check(lib().aprl_profile_query(self.handle, kernel.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def profile_reset(self) -> None:
        check(lib().aprl_profile_reset(self.handle))


class MovingWindowFields:
    """MovingWindowFields for one owner of the whole window
    (moving_window_fields.h:72-134). Reference planes are passed as arrays
    (None = the zero reference) instead of a callback."""

    def __init__(self, fields: EBFieldGrid, parameters: EBCPMLParameters):
        self.fields = fields
        p = parameters.c_struct()
        check(lib().aprl_window_install(fields.handle, C.byref(p)))

    def beginStep(self, step: int, offset: int, lower=None) -> None:
        arr, keep = plane_array(lower)
        check(lib().aprl_window_begin_step(self.fields.handle, step, offset, arr))

    def setChargeMatch(self, enable: bool = True) -> None:
        """inlet_charge_match: ChargeMatchedInlet as the new-layer completion
        with the centered normal reference (charge_matched_inlet.h:16-72)."""
        check(lib().aprl_window_set_charge_match(self.fields.handle, int(enable)))

    def beginStepCentered(self, step: int, offset: int, lower=None, ez_at_offset=None) -> None:
        arr, keep = plane_array(lower)
        ez = None if ez_at_offset is None else np.ascontiguousarray(ez_at_offset, dtype=np.float64)
        check(lib().aprl_window_begin_step_centered(self.fields.handle, step, offset, arr, dptr(ez)))

    def completeInlet(self, coupling, particles, origin) -> None:
        """completeMovingInletLayer after a shift to the window origin."""
        o = np.ascontiguousarray(origin, dtype=np.float64)
        check(lib().aprl_window_complete_inlet(self.fields.handle, coupling.handle, particles.handle, dptr(o)))

    def setReference(self, reference, translationBeta: float) -> None:
        """The reference callback on the device (aprl_window_set_reference):
        planes not passed to beginStep / setInlet / shift come from the
        BlockedCowanReference at the window's offset and step
        (movingWindowFieldValue); None restores the zero reference."""
        self._reference = reference  # kept alive while the window uses it
        check(lib().aprl_window_set_reference(self.fields.handle, None if reference is None else reference.handle,
                                              float(translationBeta)))

    def setInlet(self, ex=None, ey=None) -> None:
        ex = None if ex is None else np.ascontiguousarray(ex, dtype=np.float64)
        ey = None if ey is None else np.ascontiguousarray(ey, dtype=np.float64)
        check(lib().aprl_window_set_inlet(self.fields.handle, dptr(ex), dptr(ey)))

    def enableEnergyLedger(self, enabled: bool = True) -> None:
        """MovingWindowFields::enableEnergyLedger (moving_window_fields.h:91)."""
        check(lib().aprl_window_enable_energy_ledger(self.fields.handle, 1 if enabled else 0))

    def finishEnergyLedgerStep(self) -> None:
        """MovingWindowFields::finishEnergyLedgerStep (moving_window_fields.cpp:425-438)."""
        check(lib().aprl_window_finish_energy_step(self.fields.handle))

    def energyLedger(self) -> dict:
        """MovingWindowFieldEnergyLedger (moving_window_fields.h:39-58)."""
        out = np.zeros(13)
        check(lib().aprl_window_energy_ledger(self.fields.handle, dptr(out)))
        return {"inletOverwriteEnergyJ": out[0], "insertedEnergyJ": out[1], "croppedEnergyJ": out[2],
                "remapEnergyJ": out[3], "outwardEnergyFace": out[4:10].copy(), "overwriteEvents": int(out[10]),
                "fluxSteps": int(out[11]), "remapSteps": int(out[12])}

    def shift(self, newStep: int, newOffset: int, planes=None) -> None:
        arr, keep = plane_array(planes)
        check(lib().aprl_window_shift(self.fields.handle, newStep, newOffset, arr))
