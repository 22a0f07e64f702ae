"""Oracle reference surface current (TransverseReferenceCurrent) pinned by
closed forms.

CPU-only. transverse_reference_current.cpp cannot be built here on its own:
its prepare() needs the FFTW-based BlockedCowanReference. The restatement
is pinned instead by what the reference's code fixes: the target points
are the boundary ring of the aperture, each node once (points(), 18-26),
and a table linear in the row index is interpolated exactly, so with no
particles a coupled step leaves Jz = the table's linear function of
z = offset + k - beta (step + 1) on the ring and zero elsewhere.
"""
import numpy as np

import event_cases as E
import harness as H

BETA = 0.9375


def ring_table(points, rows, seed=3):
    rng = np.random.default_rng(seed)
    alpha = rng.normal(0, 1e3, points)
    slope = rng.normal(0, 1e2, points)
    return alpha, slope, alpha[None, :] + slope[None, :] * np.arange(rows)[:, None]


def test_reference_points_are_the_aperture_ring():
    f, c = E.oracle_case("orc")
    g = E.CELLS[:2]
    ij = c.reference_points(g)
    nx, ny, nz = E.DIMS
    x0, x1, y0, y1 = g[0], nx - g[0], g[1], ny - g[1]
    expected = {(i, j) for i in range(x0, x1 + 1) for j in (y0, y1)} | {(i, j) for j in range(y0 + 1, y1)
                                                                        for i in (x0, x1)}
    got = [tuple(v) for v in ij]
    assert len(got) == len(set(got)) == len(expected)
    assert set(got) == expected
    # the reference's order: the y faces first, then the x faces
    assert got[0] == (x0, y0) and got[1] == (x0, y1) and got[2 * (x1 - x0 + 1)] == (x0, y0 + 1)


def test_linear_table_is_interpolated_exactly():
    f, c = E.oracle_case("orc")
    c.set_particles(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0), np.zeros(0), np.zeros(0),
                    np.zeros(0, np.uint64))
    ij = c.reference_points(E.CELLS[:2])
    rows = 120
    alpha, slope, table = ring_table(len(ij), rows)
    c.set_reference_current(table, E.CELLS[:2], BETA)
    rng = np.random.default_rng(1)
    offset, step = E.OFFSET0, 0
    lower, (ex, ey) = E.planes_for(rng, range(3)), E.planes_for(rng, range(2))
    f.clear_current()
    f.window_begin(step, offset, lower)
    f.window_advance_magnetic()
    f.window_set_inlet(offset, step + 1, ex, ey)
    c.step(E.window_origin(offset), 0.0, 1)
    jz = f.view(8)
    nz = E.DIMS[2]
    expected = np.zeros_like(jz)
    for k in range(nz):
        z = float(offset) + k - BETA * (step + 1)
        expected[k, ij[:, 1], ij[:, 0]] = alpha + slope * z
    assert np.allclose(jz, expected, rtol=1e-14, atol=1e-12 * np.abs(expected).max())
    assert np.count_nonzero(jz) == nz * len(ij)
