"""MI355X (gfx950) implementation of APRL's per-timestep electromagnetic PIC
loop, driven through the C-ABI of include/aprl_gpu.h."""
from ._capi import AprlError, InvalidArgument, LogicError, OutOfRange, DeviceError, lib  # noqa: F401
from .fields import (EBCPMLParameters, EBFieldGrid, EBGridGeometry, MovingWindowFields,  # noqa: F401
                     component_shape)
from .coupling import (ConservativeSourceCoupling, EBGaussFieldInitializer, LabFieldPlane,  # noqa: F401
                       ParticleStore)
from .injection import InjectionPlanes, LabPlaneBunch, frozenLabPlaneInjection  # noqa: F401
from .reference import BlockedCowanReference  # noqa: F401
from .tuning import autotuneDecisions, setAutotune  # noqa: F401

__all__ = ["EBFieldGrid", "EBGridGeometry", "EBCPMLParameters", "MovingWindowFields", "AprlError", "lib"]
