"""The energy ledger of the moving window and the particle sums: the
oracle's restatement against the reference build (moving_window_fields.cpp:
207-218, 264-288, 387-438, 443-517 compiled where it lies) and against the
closed forms the definitions fix."""
import numpy as np
import pytest

import harness as H
from aprl_amd import bunch
from test_oracle_particles import make_window_case

needs_ref = pytest.mark.skipif(not H.have_ref(), reason="reference build not available")


def run_ledger(f, c, dims, steps, seed, finish=True):
    rng = np.random.default_rng(seed)
    offset = 50
    origin = np.zeros(3)

    def planes(comps):
        return [rng.standard_normal(int(np.prod(H.component_shape(*dims, a)[1:]))) * (10.0 if a < 3 else 10.0 / H.C_SI)
                for a in comps]

    for n in range(steps):
        lower, (ex, ey), shift = planes(range(3)), planes(range(2)), planes(range(6))
        f.clear_current()
        f.window_begin(n, offset, lower)
        f.window_advance_magnetic()
        f.window_set_inlet(offset, n + 1, ex, ey)
        c.step(origin, n * 1e-15)
        if finish:
            f.finish_ledger_step()
        offset -= 1
        f.window_shift(n + 1, offset, shift)
    return f.ledger()


@needs_ref
@pytest.mark.parametrize("shape,smooth", [("binomial3", True), ("quintic", False)])
def test_window_ledger_matches_reference_build(shape, smooth):
    dims, dz, cells = (20, 20, 40), 1.0e-6, (4, 4, 0)
    fo, co = make_window_case("orc", dims, dz, cells, 48, 17, shape, smooth)
    fr, cr = make_window_case("ref", dims, dz, cells, 48, 17, shape, smooth)
    fo.enable_ledger()
    fr.enable_ledger()
    lo = run_ledger(fo, co, dims, 4, 6)
    lr = run_ledger(fr, cr, dims, 4, 6)
    print(lo)
    assert list(lo[10:]) == [4, 4, 4]
    assert np.all(np.abs(lo[:10]) > 0)
    assert np.array_equal(lo, lr), (lo - lr)


@needs_ref
def test_window_ledger_phase_errors_match_reference_build():
    dims, dz, cells = (20, 20, 40), 1.0e-6, (4, 4, 0)
    for prefix in ("orc", "ref"):
        f, c = make_window_case(prefix, dims, dz, cells, 8, 17)
        f.enable_ledger()
        with pytest.raises(H.CpuError) as e:
            run_ledger(f, c, dims, 1, 2, finish=False)
        assert "must finish after MPI and before shift" in str(e.value)
        f2, _ = make_window_case(prefix, dims, dz, cells, 8, 17)
        z = [np.zeros(int(np.prod(H.component_shape(*dims, a)[1:]))) for a in range(3)]
        f2.window_begin(0, 50, z)
        with pytest.raises(H.CpuError) as e:
            f2.enable_ledger()
        assert "before the first step" in str(e.value)


def test_window_ledger_closed_form_flux_of_static_fields():
    """With J = 0 and fields that do not change over the step (E = 0, a
    uniform B), the Poynting vector vanishes: every face flux stays 0, the
    inlet overwrite with zero planes adds nothing, and the shift inserts
    the zero plane's energy of a uniform Bz exactly: |B|^2/(2 mu0) dV per
    aperture cell."""
    dims, dz, cells = (12, 12, 16), 1.0e-6, (3, 3, 0)
    dx = 2.5 * dz
    f = H.CpuFields("orc", *dims, dx, dx, dz, dz / H.C_SI, 1)
    f.install_window(cells)
    f.enable_ledger()
    bz = f.view(5)
    bz[...] = 1e-3
    z = [np.zeros(int(np.prod(H.component_shape(*dims, a)[1:]))) for a in range(6)]
    zb = [np.zeros(int(np.prod(H.component_shape(*dims, a)[1:]))) for a in range(6)]
    zb[5][...] = 1e-3
    f.window_begin(0, 50, z[:3])
    f.window_advance_magnetic()
    f.window_set_inlet(50, 1, z[0], z[1])
    f.advance_electric()
    f.finish_ledger_step()
    f.window_shift(1, 49, zb)
    led = f.ledger()
    assert np.all(led[4:10] == 0.0), led
    assert led[0] == 0.0
    aperture = (dims[0] - 2 * cells[0]) * (dims[1] - 2 * cells[1])
    expect = aperture * 0.5 * (1e-3) ** 2 / H.MU0 * dx * dx * dz
    assert abs(led[1] - expect) <= 1e-14 * expect, (led[1], expect)
    assert abs(led[2] - expect) <= 1e-14 * expect  # the cropped top plane holds the same uniform field
    assert abs(led[3]) <= 1e-14 * expect


@needs_ref
def test_particle_energy_sums_match_reference_build_and_closed_forms():
    rng = np.random.default_rng(4)
    dims, dz = (8, 8, 8), 1e-6
    for gamma_b in (1.0, 5.0):
        sums = {}
        for prefix in ("orc", "ref"):
            f = H.CpuFields(prefix, *dims, dz, dz, dz, dz / H.C_SI, 1)
            c = H.CpuCoupling(f)
            c.set_frame(gamma_b, 0.0, 0.3, 1e-4)
            n = 500
            x = rng.uniform(0, 8e-6, (n, 3)) if prefix == "orc" else x
            u = rng.normal(0, 3.0, (n, 3)) if prefix == "orc" else u
            m = np.full(n, bunch.ME * 7.0)
            c.set_particles(x, u, np.full(n, -bunch.QE * 7.0), m, np.ones(n), np.arange(1, n + 1))
            sums[prefix] = c.particle_energy(2e-14)
        assert np.array_equal(sums["orc"], sums["ref"]), sums["orc"] - sums["ref"]
        s = sums["orc"]
        rest = bunch.ME * 7.0 * H.C_SI ** 2
        g = np.sqrt(1.0 + np.sum(u * u, axis=1))
        assert abs(s[2] - 7.0 * n) <= 1e-12 * 7.0 * n
        assert abs(s[1] - rest * g.sum()) <= 1e-13 * s[1]
        assert abs(s[0] - rest * (g - 1.0).sum()) <= 1e-12 * s[0]
        assert abs(s[7] - rest * u[:, 2].sum()) <= 1e-12 * rest * np.abs(u[:, 2]).sum()
        # the lab frame: gamma_lab = gamma_b (gamma + beta_b u_z)
        bb = np.sqrt(1.0 - 1.0 / gamma_b ** 2)
        g_lab = gamma_b * (g + bb * u[:, 2])
        assert abs(s[3] - 7.0 * g_lab.sum()) <= 1e-12 * s[3]
        assert abs(s[6] - rest * (g_lab - 1.0).sum()) <= 1e-11 * s[6]
        z_lab = 0.3 + gamma_b * bb * H.C_SI * 2e-14 + gamma_b * (x[:, 2] - 1e-4)
        assert abs(s[5] - 7.0 * z_lab.sum()) <= 1e-12 * 7.0 * np.abs(z_lab).sum()
