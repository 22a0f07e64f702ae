"""ctypes binding of ``libaprl_gpu.so`` (the C-ABI in include/aprl_gpu.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libaprl_gpu.so")

_dp = C.POINTER(C.c_double)
_sz = C.c_size_t


class GridGeometry(C.Structure):
    """aprl_grid_geometry (EBGridGeometry, eb_field.h:54-70)."""
    _fields_ = [("nx", _sz), ("ny", _sz), ("nz", _sz), ("dx", C.c_double), ("dy", C.c_double),
                ("dz", C.c_double), ("dt", C.c_double), ("solver", C.c_int)]


class CpmlParameters(C.Structure):
    """aprl_cpml_parameters (EBCPMLParameters, eb_cpml.h:12-24)."""
    _fields_ = [("cells", _sz * 3), ("polynomial_order", C.c_double), ("target_reflection", C.c_double),
                ("kappa_max", C.c_double), ("alpha_fraction", C.c_double)]


class Magnet(C.Structure):
    """aprl_magnet (SIMagneticElement, eb_sources.h:107-144)."""
    _fields_ = [("type", C.c_int), ("center", C.c_double * 3), ("length", C.c_double), ("period", C.c_double),
                ("peak_field", C.c_double), ("polarization_angle", C.c_double), ("gaussian_fringe", C.c_int),
                ("fringe_relative_cutoff", C.c_double), ("quadrupole_gradient", C.c_double),
                ("end_ramp_periods", C.c_double), ("undulator_phase", C.c_double)]


class Boost(C.Structure):
    """aprl_boost (BoostFrameTransform::setOriginsFromGamma)."""
    _fields_ = [("gamma", C.c_double), ("lab_time_origin", C.c_double), ("lab_z_origin", C.c_double),
                ("box_z_origin", C.c_double)]


class StepReport(C.Structure):
    _fields_ = [("iterations", C.c_uint), ("force", C.c_double), ("work", C.c_double), ("tiles", C.c_uint),
                ("chunks", C.c_uint), ("active_tiles", C.c_uint)]


class GaussReport(C.Structure):
    """aprl_gauss_report (EBGaussInitializationReport, eb_initial_field.h:17-35)."""
    _fields_ = [("iterations", C.c_uint64), ("relative_residual", C.c_double), ("maximum_gauss_residual", C.c_double),
                ("total_charge", C.c_double), ("mean_beta_z", C.c_double), ("mean_beta_transverse", C.c_double),
                ("initial_electric_energy", C.c_double), ("initial_magnetic_energy", C.c_double),
                ("rms_position_cells", C.c_double * 3), ("centre_padding_rms", C.c_double * 3),
                ("static_field_guard_cells", C.c_uint64 * 3), ("transverse_poisson_padding_cells", C.c_uint64)]


class LabPlaneProjection(C.Structure):
    """aprl_lab_plane_projection (SILabPlaneProjectionReport, eb_bunch.h:36-47)."""
    _fields_ = [("input_plane_lab", C.c_double), ("reference_position_lab", C.c_double),
                ("minimum_forward_distance", C.c_double), ("maximum_forward_distance", C.c_double),
                ("recommended_minimum_reference_position", C.c_double), ("reference_time_offset_lab", C.c_double),
                ("mean_longitudinal_beta", C.c_double), ("particles", C.c_uint64)]


class BunchPlacement(C.Structure):
    """aprl_bunch_placement (SIBunchPlacementReport, eb_bunch.h:23-34)."""
    _fields_ = [("relative_tail_lab", C.c_double), ("relative_head_lab", C.c_double),
                ("reference_position_lab", C.c_double), ("head_after_lab", C.c_double),
                ("actual_head_distance", C.c_double), ("recommended_maximum_reference_position", C.c_double),
                ("particles", C.c_uint64)]


class BunchBoostReport(C.Structure):
    """aprl_bunch_boost_report (SIBunchBoostReport, eb_bunch.h:49-60)."""
    _fields_ = [("earliest_lab_event_time", C.c_double), ("latest_lab_event_time", C.c_double),
                ("maximum_absolute_drift_time", C.c_double), ("minimum_lab_gamma", C.c_double),
                ("maximum_lab_gamma", C.c_double), ("maximum_relative_momentum_round_trip_error", C.c_double),
                ("maximum_relative_gamma_round_trip_error", C.c_double)]


class InletReferenceReport(C.Structure):
    """aprl_inlet_reference_report (BlockedCowanReferenceReport, blocked_cowan_reference.h:16-56)."""
    _fields_ = [("minimum_mode_eigenvalue", C.c_double), ("minimum_pivot", C.c_double),
                ("source_charge", C.c_double), ("deposited_charge", C.c_double), ("interior_nodes", C.c_uint64 * 3)]


class InstanceStep(C.Structure):
    """aprl_instance_step: one instance of a batched window step."""
    _fields_ = [("fields", C.c_void_p), ("coupling", C.c_void_p), ("particles", C.c_void_p), ("step", C.c_size_t),
                ("offset", C.c_ssize_t), ("origin", C.c_double * 3), ("time", C.c_double), ("substeps", C.c_uint),
                ("report", StepReport)]


class InstanceRun(C.Structure):
    """aprl_instance_run: one instance of a run of batched window steps."""
    _fields_ = [("state", InstanceStep), ("reference_z", C.c_double), ("queue", C.c_void_p),
                ("look_ahead", C.c_uint64), ("plane", C.c_void_p), ("iterations", C.c_uint64),
                ("samples", C.c_uint64), ("entry_steps", C.c_uint64)]


class GaussianBeam(C.Structure):
    """aprl_gaussian_beam (generated-gaussian beam, particle_initializer.cpp:60-100)."""
    _fields_ = [("center", C.c_double * 3), ("sigma_position", C.c_double * 3), ("mean_u", C.c_double * 3),
                ("sigma_u", C.c_double * 3), ("electrons", C.c_double), ("seed", C.c_uint64), ("first", C.c_uint64),
                ("total", C.c_uint64), ("boost_gamma", C.c_double)]


class PendingGeometry(C.Structure):
    """aprl_pending_geometry (PendingWindowGeometry, pending_particle_queue.h:16-23)."""
    _fields_ = [("reference_origin_z", C.c_double), ("initial_offset", C.c_int64), ("dz", C.c_double),
                ("held_distance_per_beta", C.c_double), ("maximum_step", C.c_uint64)]


class AprlError(RuntimeError):
    """A failed C-ABI call; ``status`` is the aprl_status code."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


class InvalidArgument(AprlError, ValueError):
    pass


class OutOfRange(AprlError, IndexError):
    pass


class LogicError(AprlError):
    pass


class DeviceError(AprlError):
    pass


_STATUS = {1: InvalidArgument, 2: OutOfRange, 3: AprlError, 4: LogicError, 5: DeviceError}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m aprl_amd.build` "
                          "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.aprl_last_error.restype = C.c_char_p
    L.aprl_fields_create.argtypes = [C.POINTER(GridGeometry), C.c_int, C.POINTER(vp)]
    L.aprl_fields_destroy.argtypes = [vp]
    L.aprl_fields_destroy.restype = None
    L.aprl_fields_upload.argtypes = [vp, C.c_int, _dp, _sz]
    L.aprl_fields_download.argtypes = [vp, C.c_int, _dp, _sz]
    L.aprl_fields_download_plane.argtypes = [vp, C.c_int, _sz, _dp, _sz]
    L.aprl_fields_component_size.argtypes = [vp, C.c_int]
    L.aprl_fields_component_size.restype = _sz
    L.aprl_fields_set_cpml.argtypes = [vp, C.POINTER(CpmlParameters), _sz, _sz, C.c_int, C.c_int]
    L.aprl_fields_history_size.argtypes = [vp, _sz]
    L.aprl_fields_history_size.restype = _sz
    L.aprl_fields_history_download.argtypes = [vp, _sz, _dp, _sz]
    L.aprl_fields_history_upload.argtypes = [vp, _sz, _dp, _sz]
    for name in ("aprl_fields_clear_current", "aprl_fields_advance_electric", "aprl_fields_synchronize",
                 "aprl_profile_reset"):
        getattr(L, name).argtypes = [vp]
    L.aprl_fields_advance_magnetic.argtypes = [vp, C.POINTER(_dp)]
    L.aprl_window_install.argtypes = [vp, C.POINTER(CpmlParameters)]
    L.aprl_window_begin_step.argtypes = [vp, _sz, C.c_ssize_t, C.POINTER(_dp)]
    L.aprl_window_set_inlet.argtypes = [vp, _dp, _dp]
    L.aprl_window_shift.argtypes = [vp, _sz, C.c_ssize_t, C.POINTER(_dp)]
    L.aprl_profile_enable.argtypes = [vp, C.c_int]
    L.aprl_profile_query.argtypes = [vp, C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]
    u64p = C.POINTER(C.c_uint64)
    L.aprl_particles_create.argtypes = [vp, _sz, C.POINTER(vp)]
    L.aprl_particles_destroy.argtypes = [vp]
    L.aprl_particles_destroy.restype = None
    L.aprl_particles_upload.argtypes = [vp, _sz, _dp, _dp, _dp, _dp, _dp, u64p]
    L.aprl_particles_generate_gaussian.argtypes = [vp, C.POINTER(GaussianBeam), _sz]
    L.aprl_coupling_particle_energy.argtypes = [vp, vp, C.c_double, _dp]
    L.aprl_window_enable_energy_ledger.argtypes = [vp, C.c_int]
    L.aprl_window_finish_energy_step.argtypes = [vp]
    L.aprl_window_energy_ledger.argtypes = [vp, _dp]
    L.aprl_particles_download.argtypes = [vp, _dp, _dp, _dp, _dp, _dp, u64p]
    L.aprl_particles_count.argtypes = [vp]
    L.aprl_particles_count.restype = _sz
    L.aprl_particles_download_states.argtypes = [vp, C.POINTER(C.c_int)]
    L.aprl_particles_set_held.argtypes = [vp, C.POINTER(C.c_int), _dp]
    L.aprl_particles_download_held.argtypes = [vp, _dp]
    L.aprl_coupling_set_activation.argtypes = [vp, C.c_double]
    L.aprl_coupling_set_hold_transverse.argtypes = [vp, C.c_int]
    L.aprl_repeat_held_addition.argtypes = [C.c_double, C.c_double, C.c_uint64, _dp]
    L.aprl_pending_create.argtypes = [C.POINTER(PendingGeometry), C.c_uint64, C.POINTER(vp)]
    L.aprl_pending_destroy.argtypes = [vp]
    L.aprl_pending_destroy.restype = None
    L.aprl_pending_extract.argtypes = [vp, _sz, _dp, _dp, _dp, _dp, _dp, u64p, C.POINTER(C.c_int), _dp,
                                       C.POINTER(_sz)]
    L.aprl_pending_admit.argtypes = [vp, C.c_uint64, C.c_uint64, vp, C.POINTER(_sz)]
    L.aprl_pending_view.argtypes = [vp, C.c_uint64, C.c_uint64, _sz, _dp, u64p, C.POINTER(_sz)]
    L.aprl_pending_state.argtypes = [vp, u64p, u64p, u64p]
    L.aprl_pending_origin.argtypes = [vp, C.c_uint64, _dp]
    L.aprl_window_set_charge_match.argtypes = [vp, C.c_int]
    L.aprl_window_begin_step_centered.argtypes = [vp, _sz, C.c_ssize_t, C.POINTER(_dp), _dp]
    L.aprl_window_complete_inlet.argtypes = [vp, vp, vp, _dp]
    L.aprl_coupling_set_reference_current.argtypes = [vp, _dp, _sz, _sz, C.POINTER(_sz), C.c_double]
    L.aprl_coupling_boundary_counters.argtypes = [vp, u64p, _dp]
    L.aprl_coupling_create.argtypes = [vp, C.c_int, C.c_int, C.POINTER(vp)]
    L.aprl_coupling_destroy.argtypes = [vp]
    L.aprl_coupling_destroy.restype = None
    L.aprl_coupling_set_frame.argtypes = [vp, C.POINTER(Boost)]
    L.aprl_coupling_add_magnet.argtypes = [vp, C.POINTER(Magnet)]
    L.aprl_coupling_set_particle_boundary.argtypes = [vp, C.POINTER(_sz), C.c_double]
    L.aprl_coupling_step.argtypes = [vp, vp, _dp, C.c_double, C.c_uint, C.POINTER(StepReport)]
    L.aprl_coupling_set_split_trials.argtypes = [vp, C.c_uint, C.c_int]
    L.aprl_coupling_set_trial_batch.argtypes = [vp, C.c_uint]
    L.aprl_coupling_prepare_step.argtypes = [vp, vp, _dp]
    L.aprl_fields_initialize_gauss.argtypes = [vp, vp, _dp, C.c_int, C.c_double, C.POINTER(_sz), C.c_double, _sz,
                                               C.c_int, C.POINTER(GaussReport)]
    L.aprl_lab_plane_create.argtypes = [vp, C.c_double, C.c_double, _sz, C.POINTER(vp)]
    L.aprl_lab_plane_destroy.argtypes = [vp]
    L.aprl_lab_plane_destroy.restype = None
    L.aprl_lab_plane_sample.argtypes = [vp, C.c_double, _dp, C.POINTER(C.c_int)]
    L.aprl_lab_plane_fetch.argtypes = [vp, _sz, _dp, _dp, C.POINTER(_sz)]
    L.aprl_window_step_batch.argtypes = [C.POINTER(InstanceStep), _sz]
    L.aprl_window_run_batch.argtypes = [C.POINTER(InstanceRun), _sz, _sz]
    L.aprl_autotune.argtypes = [C.c_int]
    L.aprl_autotune_report.argtypes = [C.c_char_p, _sz, C.POINTER(_sz)]
    L.aprl_inlet_reference_create.argtypes = [C.POINTER(GridGeometry), _dp, _sz, _dp, _dp, C.c_double,
                                              C.POINTER(_sz), C.c_int, C.c_int, C.POINTER(vp),
                                              C.POINTER(InletReferenceReport)]
    L.aprl_inlet_reference_destroy.argtypes = [vp]
    L.aprl_inlet_reference_destroy.restype = None
    L.aprl_inlet_reference_download.argtypes = [vp, C.c_int, _dp, _sz]
    L.aprl_inlet_reference_plane.argtypes = [vp, C.c_int, C.c_ssize_t, _sz, C.c_double, _dp, _sz]
    L.aprl_window_set_reference.argtypes = [vp, vp, C.c_double]
    L.aprl_bunch_create.argtypes = [_sz, C.POINTER(vp)]
    L.aprl_bunch_destroy.argtypes = [vp]
    L.aprl_bunch_destroy.restype = None
    L.aprl_bunch_count.argtypes = [vp]
    L.aprl_bunch_count.restype = _sz
    L.aprl_bunch_upload.argtypes = [vp, _sz, _dp, _dp, _dp, _dp, _dp, u64p]
    L.aprl_bunch_download.argtypes = [vp, _dp, _dp, _dp, _dp, _dp, u64p, _dp]
    L.aprl_bunch_generate_plane_events.argtypes = [vp, C.POINTER(GaussianBeam), _sz]
    L.aprl_bunch_project_lab_plane_events.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_int,
                                                      C.POINTER(LabPlaneProjection)]
    L.aprl_bunch_place_relative.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double,
                                            C.POINTER(BunchPlacement)]
    L.aprl_bunch_frozen_frame_origin.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                                 C.POINTER(Boost)]
    L.aprl_bunch_boost_to_box_time_zero.argtypes = [vp, C.POINTER(Boost), C.c_double, C.c_int, C.c_double,
                                                    C.POINTER(BunchBoostReport)]
    L.aprl_bunch_freeze.argtypes = [vp, C.POINTER(Boost), C.c_double, C.c_int]
    L.aprl_fields_interior_energy.argtypes = [vp, C.POINTER(_sz), C.POINTER(_sz), _dp]
    L.aprl_fields_face_power.argtypes = [vp, C.POINTER(_sz), C.POINTER(_sz), _dp]
    L.aprl_coupling_continuity.argtypes = [vp, vp, _dp]
    L.aprl_coupling_continuity_noalloc.argtypes = [vp, vp, _dp]
    L.aprl_coupling_gauss.argtypes = [vp, vp, _dp, C.POINTER(_sz), C.POINTER(_sz), _dp]
    _lib = L
    return L


def check(status: int) -> None:
    if status:
        msg = lib().aprl_last_error().decode()
        raise _STATUS.get(status, AprlError)(status, msg)


def dptr(a: np.ndarray):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_dp)


def size3(v):
    return (_sz * 3)(*[int(x) for x in v])


def plane_array(planes):
    """(double*)[] from a list of arrays/None; returns (array, keepalive)."""
    if planes is None:
        return None, None
    keep = [None if p is None else np.ascontiguousarray(p, dtype=np.float64) for p in planes]
    arr = (_dp * len(keep))(*[dptr(p) if p is not None else _dp() for p in keep])
    return arr, keep


EXPORTED = [
    "aprl_last_error", "aprl_fields_create", "aprl_fields_destroy", "aprl_fields_upload", "aprl_fields_download",
    "aprl_fields_download_plane",
    "aprl_fields_component_size", "aprl_fields_set_cpml", "aprl_fields_history_size",
    "aprl_fields_history_download", "aprl_fields_history_upload", "aprl_fields_clear_current",
    "aprl_fields_advance_magnetic", "aprl_fields_advance_electric", "aprl_fields_synchronize",
    "aprl_window_install", "aprl_window_begin_step", "aprl_window_set_inlet", "aprl_window_shift",
    "aprl_window_set_charge_match", "aprl_window_begin_step_centered", "aprl_window_complete_inlet",
    "aprl_coupling_set_reference_current",
    "aprl_profile_enable", "aprl_profile_query", "aprl_profile_reset",
    "aprl_particles_create", "aprl_particles_destroy", "aprl_particles_upload", "aprl_particles_download",
    "aprl_particles_generate_gaussian", "aprl_coupling_particle_energy", "aprl_window_enable_energy_ledger",
    "aprl_window_finish_energy_step", "aprl_window_energy_ledger",
    "aprl_particles_count", "aprl_particles_download_states", "aprl_coupling_boundary_counters",
    "aprl_particles_set_held", "aprl_particles_download_held", "aprl_coupling_set_activation", "aprl_coupling_set_hold_transverse",
    "aprl_repeat_held_addition", "aprl_pending_create", "aprl_pending_destroy", "aprl_pending_extract",
    "aprl_pending_admit", "aprl_pending_view", "aprl_pending_state", "aprl_pending_origin",
    "aprl_coupling_create", "aprl_coupling_destroy", "aprl_coupling_set_frame",
    "aprl_coupling_add_magnet", "aprl_coupling_set_particle_boundary", "aprl_coupling_step",
    "aprl_coupling_set_split_trials", "aprl_coupling_set_trial_batch", "aprl_coupling_prepare_step",
    "aprl_window_step_batch", "aprl_window_run_batch", "aprl_autotune", "aprl_autotune_report",
    "aprl_lab_plane_create", "aprl_lab_plane_destroy", "aprl_lab_plane_sample", "aprl_lab_plane_fetch",
    "aprl_fields_initialize_gauss",
    "aprl_inlet_reference_create", "aprl_inlet_reference_destroy", "aprl_inlet_reference_download",
    "aprl_inlet_reference_plane", "aprl_window_set_reference",
    "aprl_bunch_create", "aprl_bunch_destroy", "aprl_bunch_count", "aprl_bunch_upload", "aprl_bunch_download",
    "aprl_bunch_generate_plane_events", "aprl_bunch_project_lab_plane_events", "aprl_bunch_place_relative",
    "aprl_bunch_frozen_frame_origin", "aprl_bunch_boost_to_box_time_zero", "aprl_bunch_freeze",
    "aprl_fields_interior_energy", "aprl_fields_face_power", "aprl_coupling_continuity",
    "aprl_coupling_continuity_noalloc", "aprl_coupling_gauss",
]
