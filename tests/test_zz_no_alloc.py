"""The allocation-free continuity entry point (aprl_coupling_continuity_noalloc)
against the allocating one (aprl_coupling_continuity), called through the
C-ABI: equal outputs (bitwise: the same kernels in the same order) on two
calls in a row, the second on a smaller grid with fewer particles, in both
forms; a grid beyond the workspace's vertex limit is rejected."""
import numpy as np
import pytest

from aprl_amd._capi import check, dptr, lib
from test_gpu_diagnostics import coupled_case, one_step

pytestmark = pytest.mark.gpu


def both(cg, store):
    L = lib()
    ref = np.zeros(3)
    got = np.full(3, np.nan)
    check(L.aprl_coupling_continuity(cg.handle, store.handle, dptr(ref)))
    check(L.aprl_coupling_continuity_noalloc(cg.handle, store.handle, dptr(got)))
    return ref, got


@pytest.mark.parametrize("precision", ["fp64", "fp32-split"])
def test_noalloc_continuity_equals_allocating_entry_point(precision):
    cases = [((24, 24, 48), 400, "binomial3"), ((16, 16, 32), 60, "quintic")]  # second call: smaller input
    for dims, n, shape in cases:
        print(f"no-alloc continuity {precision} {shape} {dims} n={n}", flush=True)
        gpu, win, cg, store, _, _ = coupled_case(shape, 7, dims=dims, cells=(3, 3, 0), n=n, precision=precision)
        one_step(gpu, win, cg, store, np.random.default_rng(3), dims)
        ref, got = both(cg, store)
        assert ref[1] > 0 and ref[0] <= 1e-6 * ref[2], ref
        assert np.array_equal(ref, got), (ref, got)
        # a second call on the same state gives the same values again
        _, again = both(cg, store)
        assert np.array_equal(again, ref)


def test_noalloc_continuity_rejects_grids_beyond_its_workspace():
    import aprl_amd as A
    dz = 1.0e-6
    dims = (64, 64, 600)  # 65 x 65 x 601 = 2,539,225 vertices > 2,400,000
    gpu = A.EBFieldGrid(A.EBGridGeometry(*dims, 2.5 * dz, 2.5 * dz, dz, dz / 299792458.0))
    cg = A.ConservativeSourceCoupling(gpu, "linear-binomial3")
    store = A.ParticleStore(gpu, 1)
    L = lib()
    out = np.zeros(3)
    with pytest.raises(ValueError, match="at most 2400000"):
        check(L.aprl_coupling_continuity_noalloc(cg.handle, store.handle, dptr(out)))
