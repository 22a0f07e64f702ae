"""Oracle field step pinned against the reference build and closed forms.

CPU-only: no GPU is touched here.
"""
import numpy as np
import pytest

import harness as H

needs_ref = pytest.mark.skipif(not H.have_ref(), reason="reference build oracle/_ref not present")

DZ = 1.0e-6
DT = DZ / H.C_SI


def test_cowan_equal_spacing_closed_form():
    # docs/MAXWELL_COWAN_CPML.md:48-50: alpha=7/12, beta=1/12, delta=1/48.
    c = H.cowan("orc", 1.0, 1.0, 1.0)
    assert np.allclose(c[0:3], 7 / 12, rtol=0, atol=4e-16)
    assert np.allclose(c[3:6], 1 / 12, rtol=0, atol=4e-16)
    assert np.allclose(c[6:9], 1 / 48, rtol=0, atol=4e-16)


@pytest.mark.parametrize("dx,dy,dz", [(2.5, 2.5, 1.0), (4.0, 5.0, 1.0), (1.0, 3.0, 1.0), (1.3, 1.7, 0.9)])
def test_cowan_identities(dx, dy, dz):
    c = H.cowan("orc", dx, dy, dz)
    al, be, dxy, dyz, dzx, r = c[0:3], c[3:6], c[6], c[7], c[8], c[9:12]
    # normalization (eb_field.cpp:910-922)
    assert abs(al[0] + 2 * be[1] + 2 * be[2] + 4 * dyz - 1) < 2e-15
    assert abs(al[1] + 2 * be[2] + 2 * be[0] + 4 * dzx - 1) < 2e-15
    assert abs(al[2] + 2 * be[0] + 2 * be[1] + 4 * dxy - 1) < 2e-15
    # factorization (eb_field.cpp:923-941)
    f = [4 * ((be[2] + 2 * dzx) / r[2] + (be[1] + 2 * dxy) / r[1]),
         4 * ((be[0] + 2 * dxy) / r[0] + (be[2] + 2 * dyz) / r[2]),
         4 * ((be[1] + 2 * dyz) / r[1] + (be[0] + 2 * dzx) / r[0]),
         16 * (dyz / (r[1] * r[2]) + dzx / (r[2] * r[0]) + dxy / (r[0] * r[1]))]
    assert np.allclose(f, 1.0, atol=1e-12)


@needs_ref
@pytest.mark.parametrize("dx,dy,dz", [(2.5, 2.5, 1.0), (4.0, 5.0, 1.0), (1.3, 1.7, 0.9)])
def test_cowan_matches_reference_bitwise(dx, dy, dz):
    assert np.array_equal(H.cowan("orc", dx, dy, dz), H.cowan("ref", dx, dy, dz))


def test_cowan_rejects_dz_not_smallest():
    with pytest.raises(H.CpuError) as e:
        H.cowan("orc", 1.0, 1.0, 2.0)
    assert e.value.code == 1 and "smallest" in str(e.value)


def make_pair(nx, ny, nz, dx, dy, dz, dt, solver=1):
    return H.CpuFields("orc", nx, ny, nz, dx, dy, dz, dt, solver), H.CpuFields("ref", nx, ny, nz, dx, dy, dz, dt, solver)


def assert_same(a, b, histories=False):
    for c in range(9):
        va, vb = a.view(c), b.view(c)
        assert np.array_equal(va, vb), f"component {c} differs: max {np.max(np.abs(va - vb))}"
    if histories:
        for h in range(12):
            assert np.array_equal(a.history(h), b.history(h)), f"history {h} differs"


@needs_ref
@pytest.mark.parametrize("solver", [0, 1])
def test_pec_step_bitwise(solver):
    dims = (9, 8, 10)
    dx, dy, dz = 2.5 * DZ, 2.0 * DZ, DZ
    dt = DT if solver else 0.5 * DT
    o, r = make_pair(*dims, dx, dy, dz, dt, solver)
    for f in (o, r):
        H.seeded_fields(f, 7, components=range(9), scale=1e3)
    for _ in range(5):
        for f in (o, r):
            f.advance_magnetic()
            f.advance_electric()
    assert_same(o, r)


@needs_ref
def test_cowan_halo_path_bitwise():
    dims = (7, 6, 9)
    dx, dy, dz = 2.0 * DZ, 1.5 * DZ, DZ
    o, r = make_pair(*dims, dx, dy, dz, DT)
    rng = np.random.default_rng(3)
    halo = []
    for a in range(3):
        shp = H.component_shape(*dims, a)
        halo.append(rng.standard_normal(shp[1] * shp[2]))
    halo += [None, rng.standard_normal(H.component_shape(*dims, 1)[1] * H.component_shape(*dims, 1)[2]), None]
    for f in (o, r):
        H.seeded_fields(f, 11, scale=1e3)
        f.advance_magnetic(halo)
    assert_same(o, r)


@needs_ref
@pytest.mark.parametrize("lower_z,upper_z,cells", [(True, True, (3, 4, 2)), (False, False, (3, 3, 0)),
                                                   (True, False, (2, 3, 0))])
def test_cpml_step_bitwise(lower_z, upper_z, cells):
    dims = (12, 11, 16)
    dx, dy, dz = 2.0 * DZ, 2.5 * DZ, DZ
    o, r = make_pair(*dims, dx, dy, dz, DT)
    for f in (o, r):
        f.install_cpml(cells, order=3.0, R=1e-8, kappa=8.0, alpha=0.05, lower_z=lower_z, upper_z=upper_z)
        H.seeded_fields(f, 5, components=range(9), scale=1e3)
    for _ in range(6):
        for f in (o, r):
            f.advance_magnetic()
            f.advance_electric()
    assert_same(o, r, histories=True)


@needs_ref
def test_cpml_rejects_z_slab_on_open_face():
    for p in ("orc", "ref"):
        f = H.CpuFields(p, 12, 11, 16, 2 * DZ, 2 * DZ, DZ, DT)
        with pytest.raises(H.CpuError) as e:
            f.install_cpml((2, 3, 3), lower_z=True, upper_z=False)
        assert e.value.code == 1 and "upper z CPML" in str(e.value)


def _planes(rng, dims, comps):
    out = []
    for a in comps:
        shp = H.component_shape(*dims, a)
        out.append(rng.standard_normal(shp[1] * shp[2]) * (1.0 if a < 3 else 1.0 / H.C_SI))
    return out


def run_window(f, dims, steps, seed, preview=True):
    """Moving-window protocol of Simulation::solve (simulation.cpp:518-583)
    with the conservative-source preview/undo/commit sequence
    (simulation_source_coupling.cpp:16-65) for J supplied by the caller."""
    rng = np.random.default_rng(seed)
    offset = 40
    for n in range(steps):
        f.clear_current()
        f.window_begin(n, offset, _planes(rng, dims, range(3)))
        f.window_advance_magnetic()
        ex, ey = _planes(rng, dims, range(2))
        f.window_set_inlet(offset, n + 1, ex, ey)
        if preview:
            old = [f.view(a).copy() for a in range(3)]
            f.window_preview(True)
            f.advance_electric()
            f.window_preview(False)
            star = [f.view(a).copy() for a in range(3)]
            # swap: the grid holds E^n again, then E* returns for the commit
            for a in range(3):
                f.view(a)[...] = star[a]
            f.window_undo_preview()
            js = _planes(rng, dims, range(3))  # a thin current slab near the middle
            for a in range(3):
                j = f.view(6 + a)
                j[dims[2] // 2, :, :] = js[a].reshape(j.shape[1], j.shape[2]) * 1e-3
                e = f.view(a)
                mask = np.zeros(e.shape, bool)
                if a == 0:
                    mask[:, 1:dims[1], :] = True
                    mask[0, :, :] = False
                    mask[dims[2], :, :] = False
                elif a == 1:
                    mask[:, :, 1:dims[0]] = True
                    mask[0, :, :] = False
                    mask[dims[2], :, :] = False
                else:
                    mask[:, 1:dims[1], 1:dims[0]] = True
                nz = mask & (j != 0)
                e[nz] += (-1.0 * (f_dt := DT)) / H.EPS0 * j[nz]
            f.window_commit()
            del old
        else:
            f.advance_electric()
        offset -= 1
        f.window_shift(n + 1, offset, _planes(rng, dims, range(6)))


@needs_ref
@pytest.mark.parametrize("preview", [False, True])
def test_window_steps_bitwise(preview):
    dims = (14, 13, 20)
    dx, dy, dz = 2.0 * DZ, 2.0 * DZ, DZ
    o, r = make_pair(*dims, dx, dy, dz, DT)
    for f in (o, r):
        f.install_window((3, 3, 0), alpha=0.0)
        H.seeded_fields(f, 9, scale=1e3)
    for f in (o, r):
        run_window(f, dims, 5, 123, preview=preview)
    assert_same(o, r, histories=True)


@needs_ref
def test_energy_and_poynting_cells_bitwise():
    dims = (6, 5, 7)
    o, r = make_pair(*dims, 2 * DZ, 2 * DZ, DZ, DT)
    for f in (o, r):
        H.seeded_fields(f, 2, scale=1e4)
    for (i, j, k) in [(0, 0, 0), (2, 3, 4), (5, 4, 6)]:
        assert o.energy_density(i, j, k) == r.energy_density(i, j, k)
        for ax in range(3):
            assert o.poynting(i, j, k, ax) == r.poynting(i, j, k, ax)


def test_cowan_axial_plane_wave_translates_exactly():
    """core_checks.cpp:771-823: with c*dt=dz an axial wave moves one cell per
    step with error < 2e-11 after 48 steps (NoBoundary: interior only)."""
    nz, ny, dz = 256, 128, 1.0e-5
    f = H.CpuFields("orc", 4, ny, nz, 4 * dz, 5 * dz, dz, dz / H.C_SI, 1)
    theta = 2 * np.pi / 16
    ex, by = f.view(0), f.view(4)
    k = np.arange(nz + 1, dtype=np.float64)
    ex[...] = np.sin(theta * k)[:, None, None]
    by[...] = (np.sin(theta * (np.arange(nz) + 1.0)) / H.C_SI)[:, None, None]
    # The reference check removes the boundary (NoBoundary); the PEC faces
    # here do not reach the sampled column k in [52, 204) at j=ny/2 within 48
    # steps because the wave is transversely uniform (PEC parity keeps it so).
    for _ in range(48):
        f.advance_magnetic()
        f.advance_electric()
    ks = np.arange(48 + 4, nz - 48 - 4)
    err = np.max(np.abs(f.view(0)[ks, ny // 2, 1] - np.sin(theta * (ks - 48.0))))
    assert err < 2e-11, err
