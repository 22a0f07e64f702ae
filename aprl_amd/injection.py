"""The laboratory-plane input route of a moving-window run.

``LabPlaneBunch`` is a device-resident bunch in laboratory coordinates
(aprl_bunch_*) with SIBunchPreprocessor's functions (eb_bunch.h:62-109) as
methods; ``frozenLabPlaneInjection`` runs Simulation's startup sequence for
HDF5 version-4 plane-arrival records with frozen injection and the
transverse hold (simulation.cpp:1983-2245): projection to the common lab
time, relative placement with the fit check against the first magnetic
interaction entrance, the frame origin at the earliest activation-plane
crossing, the boost to box time zero with its round-trip limit, and the
frozen flags.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._capi import (Boost, BunchBoostReport, BunchPlacement, GaussianBeam, LabPlaneProjection, check, dptr, lib)

C_SI = 299792458.0


def _report(struct) -> dict:
    return {name: getattr(struct, name) for name, _ in struct._fields_}


class LabPlaneBunch:
    """aprl_bunch: x, u, charge, mass, weight, id and the held speed of
    every macro-particle on the device."""

    def __init__(self, capacity: int = 0):
        h = C.c_void_p()
        check(lib().aprl_bunch_create(int(capacity), C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().aprl_bunch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return int(lib().aprl_bunch_count(self.handle))

    def upload(self, position, u, charge, mass, weight=None, ids=None):
        n = len(charge)
        x = np.ascontiguousarray(position, dtype=np.float64).reshape(n, 3)
        uu = np.ascontiguousarray(u, dtype=np.float64).reshape(n, 3)
        q = np.ascontiguousarray(charge, dtype=np.float64)
        m = np.ascontiguousarray(mass, dtype=np.float64)
        w = None if weight is None else np.ascontiguousarray(weight, dtype=np.float64)
        i = np.ascontiguousarray(np.arange(1, n + 1) if ids is None else ids, dtype=np.uint64)
        check(lib().aprl_bunch_upload(self.handle, n, dptr(x), dptr(uu), dptr(q), dptr(m), dptr(w),
                                      i.ctypes.data_as(C.POINTER(C.c_uint64))))

    def download(self):
        """(position (n,3), u (n,3), charge, mass, weight, ids, held beta)."""
        n = len(self)
        x, u = np.zeros((n, 3)), np.zeros((n, 3))
        q, m, w, beta = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
        ids = np.zeros(n, dtype=np.uint64)
        check(lib().aprl_bunch_download(self.handle, dptr(x), dptr(u), dptr(q), dptr(m), dptr(w),
                                        ids.ctypes.data_as(C.POINTER(C.c_uint64)), dptr(beta)))
        return x, u, q, m, w, ids, beta

    def generatePlaneEvents(self, count, center_xy, sigma_xy, sigma_time, mean_u, sigma_u, electrons, seed,
                            first=0, total=None, mean_time=0.0):
        """Plane-arrival records of the generated-gaussian beam
        (aprl_bunch_generate_plane_events): axis 2 is the arrival time
        offset in seconds."""
        b = GaussianBeam()
        center = [center_xy[0], center_xy[1], mean_time]
        sigma = [sigma_xy[0], sigma_xy[1], sigma_time]
        for a in range(3):
            b.center[a], b.sigma_position[a] = center[a], sigma[a]
            b.mean_u[a], b.sigma_u[a] = mean_u[a], sigma_u[a]
        b.electrons, b.seed, b.first = float(electrons), int(seed), int(first)
        b.total = int(count if total is None else total)
        b.boost_gamma = 1.0
        check(lib().aprl_bunch_generate_plane_events(self.handle, C.byref(b), int(count)))

    # SIBunchPreprocessor (eb_bunch.h:62-109)
    def projectLabPlaneEventsToSnapshot(self, inputPlaneLab, referencePositionLab, longitudinalOffsetLab=0.0,
                                        holdTransverseUntilActivation=False) -> dict:
        r = LabPlaneProjection()
        check(lib().aprl_bunch_project_lab_plane_events(self.handle, float(inputPlaneLab),
                                                        float(referencePositionLab), float(longitudinalOffsetLab),
                                                        int(bool(holdTransverseUntilActivation)), C.byref(r)))
        return _report(r)

    def placeRelativeLabSnapshot(self, referencePositionLab, firstInteractionEntranceLab=None,
                                 recommendationMarginLab=0.0) -> dict:
        r = BunchPlacement()
        has = firstInteractionEntranceLab is not None
        check(lib().aprl_bunch_place_relative(self.handle, int(has),
                                              float(firstInteractionEntranceLab) if has else float("nan"),
                                              float(referencePositionLab), float(recommendationMarginLab),
                                              C.byref(r)))
        return _report(r)

    def frozenFrameOrigin(self, gammaBoost, inputPlaneLab, activationPlaneLab, dz,
                          holdTransverseUntilActivation=True):
        """(gamma, lab time origin, lab z origin, box z origin)."""
        f = Boost()
        check(lib().aprl_bunch_frozen_frame_origin(self.handle, float(gammaBoost), float(inputPlaneLab),
                                                   float(activationPlaneLab), float(dz),
                                                   int(bool(holdTransverseUntilActivation)), C.byref(f)))
        return (f.gamma, f.lab_time_origin, f.lab_z_origin, f.box_z_origin)

    def boostLabSnapshotToBoxTimeZero(self, frame, snapshotTimeLab=0.0, holdTransverseUntilActivation=False,
                                      inputPlaneLab=0.0) -> dict:
        f = Boost(*[float(v) for v in frame])
        r = BunchBoostReport()
        check(lib().aprl_bunch_boost_to_box_time_zero(self.handle, C.byref(f), float(snapshotTimeLab),
                                                      int(bool(holdTransverseUntilActivation)), float(inputPlaneLab),
                                                      C.byref(r)))
        return _report(r)

    def freeze(self, frame, activationPlaneLab, holdTransverseUntilActivation=True):
        f = Boost(*[float(v) for v in frame])
        check(lib().aprl_bunch_freeze(self.handle, C.byref(f), float(activationPlaneLab),
                                      int(bool(holdTransverseUntilActivation))))


@dataclass
class InjectionPlanes:
    """beam.reference of a frozen lab-plane injection card (config.cpp:
    1433-1486): input plane, activation plane (at or upstream of it),
    initial centre, and the first magnetic interaction entrance (None
    without magnets)."""
    input_plane_z: float
    activation_plane_z: float
    initial_center_z: float
    first_interaction_entrance: float | None = None


def frozenLabPlaneInjection(bunch: LabPlaneBunch, planes: InjectionPlanes, gamma_boost: float, dz: float):
    """Simulation's frozen lab-plane injection with the transverse hold
    (simulation.cpp:1983-2245; the moving window requires both,
    config.cpp:2071-2076) on the device bunch. Returns (frame, reports):
    the BoostFrameTransform origins of the run and the projection,
    placement and boost reports. Raises the reference's errors."""
    from ._capi import AprlError
    if planes.activation_plane_z > planes.input_plane_z:
        raise AprlError(1, "activation_plane_z must be at or upstream of input_plane_z")
    rp = bunch.projectLabPlaneEventsToSnapshot(planes.input_plane_z, planes.initial_center_z, 0.0, True)
    margin = gamma_boost * dz  # simulation.cpp:2019-2022
    rl = bunch.placeRelativeLabSnapshot(planes.initial_center_z, planes.first_interaction_entrance, margin)
    if planes.first_interaction_entrance is not None:
        # simulation.cpp:2024-2070
        limit = planes.first_interaction_entrance - margin
        scale = max(1.0, max(abs(planes.first_interaction_entrance), abs(rl["head_after_lab"])))
        if rl["head_after_lab"] >= limit - 128.0 * np.finfo(float).eps * scale:
            raise AprlError(3, "The reconstructed common-time laboratory bunch does not fit before the first "
                               f"magnetic interaction region: bunch_head_z={rl['head_after_lab']:.10g} m, "
                               f"required_head_limit={limit:.10g} m")
    frame = bunch.frozenFrameOrigin(gamma_boost, planes.input_plane_z, planes.activation_plane_z, dz, True)
    rb = bunch.boostLabSnapshotToBoxTimeZero(frame, 0.0, True, planes.input_plane_z)
    # simulation.cpp:2316-2331
    if rb["maximum_relative_momentum_round_trip_error"] > 1e-10 or \
            rb["maximum_relative_gamma_round_trip_error"] > 1e-10:
        raise AprlError(3, "Lorentz boost round-trip error exceeds 1e-10")
    bunch.freeze(frame, planes.activation_plane_z, True)
    return frame, {"projection": rp, "placement": rl, "boost": rb}
