"""Seeded synthetic bunches.

``seeded_gaussian`` follows the generator of particle_initializer.cpp:24-107
(splitmix64 stream per particle id, Box-Muller pairs, seed ^ 0xd6e8feb86659fd93
for the momenta) in vectorised numpy; ``lab_to_box_u`` is the light-front
Lorentz transform of boostframe.cpp:110-164. numpy's log/cos may differ from
glibc by an ulp, so a bunch made here is a deterministic input, identical for
every implementation it is fed to, not a bit-copy of the reference's.
"""
from __future__ import annotations

import numpy as np

C_SI = 299792458.0
QE = 1.602176634e-19
ME = 9.1093837139e-31

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(v: np.ndarray) -> np.ndarray:
    v = np.asarray(v, dtype=np.uint64)
    with np.errstate(over="ignore"):
        v = v + _GOLDEN
        v = (v ^ (v >> np.uint64(30))) * _M1
        v = (v ^ (v >> np.uint64(27))) * _M2
    return v ^ (v >> np.uint64(31))


def uniform_open(seed: int, particle: np.ndarray, component: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        mixed = splitmix64(np.uint64(seed) ^ splitmix64(particle + np.uint64(0x100000001B3) * np.uint64(component)))
    mantissa = (mixed >> np.uint64(11)) | np.uint64(1)
    value = mantissa.astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.minimum(value, 1.0 - np.finfo(np.float64).eps)


def gaussian(seed: int, particle: np.ndarray, pair: int) -> np.ndarray:
    first = uniform_open(seed, particle, 2 * pair)
    second = uniform_open(seed, particle, 2 * pair + 1)
    return np.sqrt(-2.0 * np.log(first)) * np.cos(2.0 * np.pi * second)


def seeded_gaussian(n: int, seed: int, center, sigma_position, mean_u, sigma_u, electrons: float,
                    first_id: int = 0):
    """Positions (n,3) [m], proper velocities (n,3), charge, mass, ids.

    ``electrons`` is the total represented electron count (charge -e each).
    """
    ids = np.arange(first_id, first_id + n, dtype=np.uint64)
    x = np.empty((n, 3))
    u = np.empty((n, 3))
    for a in range(3):
        x[:, a] = center[a] + sigma_position[a] * gaussian(seed, ids, a)
        u[:, a] = mean_u[a] + sigma_u[a] * gaussian(seed ^ 0xD6E8FEB86659FD93, ids, a)
    macro = electrons / n
    q = np.full(n, -QE * macro)
    m = np.full(n, ME * macro)
    return x, u, q, m, ids + np.uint64(1)


def gamma_from_u(u: np.ndarray) -> np.ndarray:
    """BoostFrameTransform::gammaFromProperVelocity (boostframe.cpp:74-87)."""
    scale = np.maximum(1.0, np.max(np.abs(u), axis=-1))
    inv = 1.0 / scale
    return scale * np.sqrt(inv * inv + np.sum((u * inv[..., None]) ** 2, axis=-1))


def lab_to_box_u(u_lab: np.ndarray, gamma_boost: float) -> np.ndarray:
    """properVelocityLabToBox through the positive light-front component."""
    gb = np.sqrt(gamma_boost - 1.0) * np.sqrt(gamma_boost + 1.0)
    rapidity = gamma_boost + gb
    g = gamma_from_u(u_lab)
    mt2 = u_lab[:, 0] ** 2 + u_lab[:, 1] ** 2 + 1.0
    positive = np.where(u_lab[:, 2] >= 0.0, g + u_lab[:, 2], mt2 / (g - u_lab[:, 2]))
    pb = positive / rapidity
    out = u_lab.copy()
    out[:, 2] = 0.5 * (pb - mt2 / pb)
    return out
