"""Host-side mirror of the moving window's pending admission queue over the
C-ABI (PendingParticleQueue, pending_particle_queue.h:25-76).

Held particles upstream of the inlet wait in host memory and enter a device
``ParticleStore`` as held entries, their z given by the exact repeated held
addition (``repeatHeldAddition``, pending_particle_queue.h:11-14).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import check, dptr, lib
from .coupling import ParticleStore

# visitNear's look-ahead of the wide longitudinal shapes (simulation.cpp:4171)
SUPPORT_LOOK_AHEAD = 8

_u64p = C.POINTER(C.c_uint64)


def repeatHeldAddition(initial: float, increment: float, count: int) -> float:
    """The value of ``count`` separately rounded binary64 additions
    x += increment from x = initial."""
    out = C.c_double()
    check(lib().aprl_repeat_held_addition(float(initial), float(increment), int(count), C.byref(out)))
    return out.value


@dataclass
class PendingWindowGeometry:
    """PendingWindowGeometry (pending_particle_queue.h:16-23)."""
    referenceOriginZ: float
    initialOffset: int
    dz: float
    heldDistancePerBeta: float  # c * dt
    maximumStep: int


class PendingParticleQueue:
    def __init__(self, geometry: PendingWindowGeometry, initialStep: int = 0):
        g = _capi.PendingGeometry(float(geometry.referenceOriginZ), int(geometry.initialOffset), float(geometry.dz),
                                  float(geometry.heldDistancePerBeta), int(geometry.maximumStep))
        h = C.c_void_p()
        check(lib().aprl_pending_create(C.byref(g), int(initialStep), C.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().aprl_pending_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def extract(self, position, u, charge, mass, weight, ids, frozen, beta):
        """Moves the waiting held particles into the queue; returns the kept
        ones (position, u, charge, mass, weight, ids, frozen, beta) in their
        original order."""
        x = np.array(position, dtype=np.float64).reshape(-1, 3)
        n = x.shape[0]
        uu = np.array(u, dtype=np.float64).reshape(n, 3)
        q = np.array(np.broadcast_to(charge, (n,)), dtype=np.float64)
        m = np.array(np.broadcast_to(mass, (n,)), dtype=np.float64)
        w = np.array(np.broadcast_to(weight, (n,)), dtype=np.float64)
        i = np.array(ids, dtype=np.uint64)
        fr = np.array(frozen, dtype=np.int32)
        b = np.array(beta, dtype=np.float64)
        kept = C.c_size_t()
        check(lib().aprl_pending_extract(self.handle, n, dptr(x.ravel()), dptr(uu.ravel()), dptr(q), dptr(m),
                                         dptr(w), i.ctypes.data_as(_u64p), fr.ctypes.data_as(C.POINTER(C.c_int)),
                                         dptr(b), C.byref(kept)))
        k = kept.value
        return x[:k], uu[:k], q[:k], m[:k], w[:k], i[:k], fr[:k], b[:k]

    def admit(self, step: int, store: ParticleStore, lookAhead: int = 0) -> int:
        """Appends the particles due by step + lookAhead to the store as held
        entries (z at the start of ``step``); returns how many."""
        n = C.c_size_t()
        check(lib().aprl_pending_admit(self.handle, int(step), int(lookAhead), store.handle, C.byref(n)))
        return n.value

    def visitNear(self, step: int, lookAhead: int):
        """(positions (n,3), ids) of the waiting particles due by
        step + lookAhead, z at the start of ``step``."""
        n = C.c_size_t()
        check(lib().aprl_pending_view(self.handle, int(step), int(lookAhead), 0, None, None, C.byref(n)))
        x = np.empty((n.value, 3))
        i = np.empty(n.value, dtype=np.uint64)
        if n.value:
            check(lib().aprl_pending_view(self.handle, int(step), int(lookAhead), n.value, dptr(x.ravel()),
                                          i.ctypes.data_as(_u64p), C.byref(n)))
        return x, i

    def _state(self):
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(lib().aprl_pending_state(self.handle, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def size(self) -> int:
        return self._state()[0]

    def nextDueStep(self) -> int:
        return self._state()[1]

    def admittedCount(self) -> int:
        return self._state()[2]

    def origin(self, step: int) -> float:
        z = C.c_double()
        check(lib().aprl_pending_origin(self.handle, int(step), C.byref(z)))
        return z.value
