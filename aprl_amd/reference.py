"""BlockedCowanReference (blocked_cowan_reference.h:68-129) on the device:
the moving window's inlet reference prepared from the source particles
(aprl_inlet_reference_*), and movingWindowFieldValue over a plane."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._capi import InletReferenceReport, check, dptr, lib
from .fields import EBGridGeometry, component_shape

SOURCE_SHAPES = {"cic": -1, "quintic-compensated": 0, "linear-binomial3": 2}


class BlockedCowanReference:
    """The reference field of the source particles (positions (n, 3) in the
    box frame, charges) on the global grid `geometry` with origin `origin`,
    the Cowan operator at `beta` (the mean held beta, mass weighted) and
    Dirichlet `guards`; `shape` is the source deposit ("cic",
    "quintic-compensated", "linear-binomial3"), `transverseXBinomial3` its
    x weights."""

    def __init__(self, geometry: EBGridGeometry, origin, position, charge, beta: float, guards,
                 shape: str = "cic", transverseXBinomial3: bool = False):
        self.geometry = geometry
        g = geometry.c_struct()
        o = np.ascontiguousarray(origin, dtype=np.float64)
        x = np.ascontiguousarray(position, dtype=np.float64).reshape(-1, 3)
        q = np.ascontiguousarray(charge, dtype=np.float64)
        gd = (C.c_size_t * 3)(*[int(v) for v in guards])
        r = InletReferenceReport()
        h = C.c_void_p()
        check(lib().aprl_inlet_reference_create(C.byref(g), dptr(o), len(q), dptr(x.ravel()), dptr(q), float(beta),
                                                gd, SOURCE_SHAPES[shape], int(bool(transverseXBinomial3)),
                                                C.byref(h), C.byref(r)))
        self.handle = h
        self.report = {"minimumModeEigenvalue": r.minimum_mode_eigenvalue, "minimumPivot": r.minimum_pivot,
                       "sourceCharge": r.source_charge, "depositedCharge": r.deposited_charge,
                       "interiorNodes": tuple(int(v) for v in r.interior_nodes)}

    def close(self):
        if getattr(self, "handle", None):
            lib().aprl_inlet_reference_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def field(self, component: int) -> np.ndarray:
        """A whole native component (nz, ny, nx) of E(0), B(-dt/2)."""
        g = self.geometry
        shape = component_shape(g, component)
        out = np.zeros(int(np.prod(shape)))
        check(lib().aprl_inlet_reference_download(self.handle, int(component), dptr(out), out.size))
        return out.reshape(shape)

    def movingValue(self, component: int, k: int, step: int, translationBeta: float) -> np.ndarray:
        """movingWindowFieldValue over the component's x-y nodes at plane k."""
        g = self.geometry
        shape = component_shape(g, component)[1:]
        out = np.zeros(int(np.prod(shape)))
        check(lib().aprl_inlet_reference_plane(self.handle, int(component), int(k), int(step), float(translationBeta),
                                               dptr(out), out.size))
        return out.reshape(shape)
