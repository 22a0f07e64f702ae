"""The per-step tile sort on grids with many tiles.

After a coupled step the store holds the particles in the stable order of
their start-of-step tile keys (8x8x4 cell tiles, x fastest); the expected
permutation is numpy's stable argsort of the same keys. Cases: 512 tiles
(two 8-bit digit passes), 131072 tiles with 3e5 particles (three passes,
more than 2^18 particles, thousands of digit-table workgroups).
"""
import numpy as np
import pytest

import aprl_amd as A
from aprl_amd import bunch

pytestmark = pytest.mark.gpu

DZ = 1.0e-6
DX = 2.0 * DZ


def sorted_after_one_step(dims, cells, n, seed):
    g = A.EBGridGeometry(*dims, DX, DX, DZ, DZ / bunch.C_SI)
    gpu = A.EBFieldGrid(g)
    win = A.MovingWindowFields(gpu, A.EBCPMLParameters(cells))
    cg = A.ConservativeSourceCoupling(gpu, "linear-binomial3")
    cg.setFrame(5.0)
    cg.setParticleBoundary(cells)
    rng = np.random.default_rng(seed)
    lo = np.array([cells[0] + 1, cells[1] + 1, 2], float)
    hi = np.array([dims[0] - cells[0] - 1, dims[1] - cells[1] - 1, dims[2] - 2], float)
    x = (lo + rng.uniform(0.0, 1.0, (n, 3)) * (hi - lo)) * np.array([DX, DX, DZ])
    u = rng.normal(0.0, 0.05, (n, 3))
    ids = rng.permutation(n).astype(np.uint64) + np.uint64(1)
    store = A.ParticleStore(gpu, n)
    store.upload(x, u, -bunch.QE, bunch.ME, np.ones(n), ids)
    print(f"sort case {dims} n={n}", flush=True)
    gpu.clearCurrent()
    win.beginStep(0, 50)
    gpu.advanceMagnetic()
    win.setInlet()
    cg.step(store, np.zeros(3), 0.0)
    _, _, _, _, _, out = store.download()
    cx = np.clip(np.floor(x[:, 0] / DX).astype(np.int64), 0, dims[0] - 1)
    cy = np.clip(np.floor(x[:, 1] / DX).astype(np.int64), 0, dims[1] - 1)
    cz = np.clip(np.floor(x[:, 2] / DZ).astype(np.int64), 0, dims[2] - 1)
    ntx, nty = -(-dims[0] // 8), -(-dims[1] // 8)
    key = ((cz // 4) * nty + cy // 8) * ntx + cx // 8
    return out, ids[np.argsort(key, kind="stable")], key, cg


def test_sort_two_digit_passes():
    out, expected, key, cg = sorted_after_one_step((64, 64, 32), (4, 4, 0), 20000, 1)
    assert key.max() >= 256
    assert np.array_equal(out, expected)
    assert cg.last_tiles == len(np.unique(key))


def test_sort_three_digit_passes_beyond_2_18_particles():
    out, expected, key, cg = sorted_after_one_step((128, 128, 2048), (4, 4, 0), 300_000, 2)
    assert key.max() >= 65536
    assert len(expected) > 2 ** 18
    assert np.array_equal(out, expected)
    assert cg.last_tiles == len(np.unique(key))
