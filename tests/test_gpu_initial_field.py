"""The Gauss-consistent initial self-field on the GPU
(aprl_fields_initialize_gauss) against the oracle's restatement, which is
bitwise equal to the reference build's EBGaussFieldInitializer
(tests/test_oracle_initial_field.py):

* the relativistic-Poisson, electrostatic and Cowan relativistic-Poisson
  models and the padded transverse-CPML source: E and B within 1e-12
  relative L2 of the oracle's (the CG dot products are double-double sums
  in fixed order instead of serial long-double sums; the stencils, updates
  and field constructions are the reference's), the same CG iteration
  count, and the report's charge, mean velocity, moments and energies
  within 1e-12;
* the FP32-split grid: the same solve from the FP32-split store (charges,
  masses, proper velocities and positions held in FP32), its E and B
  stored in FP32: fields and report within 1e-6 of the FP64 oracle's;
* the FEL segment at full size (128x128x1024, 4e6 quiet-start
  macro-particles, resonant boost, box-frame beta_z about 0.2) with the
  Cowan model at the reference's default tolerance and an 8-cell
  transverse guard: converged, and the discrete Gauss residual of the
  stored field within 4x the tolerance (the reference's own post-solve
  check);
* the bending-radiation bunch (box-frame gamma about 10 at gamma_b = 5) is
  rejected by the Cowan operator's positive-definiteness bound with the
  reference's message, as the reference rejects it.
"""
import numpy as np
import pytest

import config_cases as CC
import harness as H
from aprl_amd import workloads as W
from test_oracle_initial_field import bunch_case, gauss

pytestmark = pytest.mark.gpu

REPORT_KEYS = ["iterations", "relativeResidual", "maximumGaussResidual", "totalCharge", "meanBetaZ",
               "meanBetaTransverse", "initialElectricEnergy", "initialMagneticEnergy"]


def gpu_case(dims, spacing, x, u, q, m, precision="fp64"):
    import aprl_amd as A
    f = A.EBFieldGrid(A.EBGridGeometry(*dims, *spacing, spacing[2] / H.C_SI), precision=precision)
    st = A.ParticleStore(f, len(q))
    st.upload(x, u, q, m, None, np.arange(1, len(q) + 1, dtype=np.uint64))
    return f, st


@pytest.mark.parametrize("model,include_cpml,precision", [
    ("relativistic-poisson", False, "fp64"), ("electrostatic-poisson", False, "fp64"),
    ("cowan-relativistic-poisson", False, "fp64"), ("cowan-relativistic-poisson", True, "fp64"),
    ("cowan-relativistic-poisson", False, "fp32-split")])
def test_initial_field_matches_oracle(model, include_cpml, precision):
    import aprl_amd as A
    dims, spacing = (16, 16, 32), (2.5e-6, 2.5e-6, 1e-6)
    x, u, q, m = bunch_case(dims, spacing, 300, 0.3)
    guard = (3, 3, 2)
    print(f"initial field {model} cpml={include_cpml} {precision}", flush=True)
    orc = H.CpuFields("orc", *dims, *spacing, spacing[2] / H.C_SI, 1)
    ro = gauss("orc", orc, np.zeros(3), x, u, q, m, model, guard, include_cpml=include_cpml)
    f, st = gpu_case(dims, spacing, x, u, q, m, precision)
    rg = A.EBGaussFieldInitializer.initialize(f, st, np.zeros(3), model, 1e-2, guard, 1e-10, 5000, include_cpml)
    print({k: rg[k] for k in REPORT_KEYS}, "oracle iterations", ro[0], flush=True)
    assert rg["iterations"] == int(ro[0])
    tol = 1e-6 if precision == "fp32-split" else 1e-12
    errs = [CC.rel_l2(f.component(a), orc.view(a)) for a in range(6)]
    print("E/B rel L2", [f"{e:.2e}" for e in errs])
    vec_e = np.sqrt(sum(np.sum((f.component(a) - orc.view(a)) ** 2) for a in range(3)) /
                    sum(np.sum(orc.view(a) ** 2) for a in range(3)))
    assert vec_e <= tol, errs
    if model != "electrostatic-poisson":
        vec_b = np.sqrt(sum(np.sum((f.component(a) - orc.view(a)) ** 2) for a in range(3, 5)) /
                        sum(np.sum(orc.view(a) ** 2) for a in range(3, 5)))
        assert vec_b <= tol, errs
    else:
        assert not np.any(f.component(3)) and not np.any(f.component(4))
    for i, k in enumerate(REPORT_KEYS[3:], start=3):
        assert rg[k] == pytest.approx(ro[i], rel=tol, abs=1e-300), k
    assert rg["relativeResidual"] <= 4e-10
    assert np.allclose(rg["rmsPositionCells"], ro[8:11], rtol=tol)
    assert np.allclose(rg["centrePaddingRms"], ro[11:14], rtol=tol)
    assert list(rg["staticFieldGuardCells"]) == list(ro[14:17]) and rg["transversePoissonPaddingCells"] == ro[17]


def test_fel_segment_cowan_initial_field_passes_gauss_check():
    import aprl_amd as A
    w = W.FEL_C3
    fields, window, coupling, store = W.setup_gpu(w, "quintic-compensated", "fp64")
    print("FEL initial field", flush=True)
    r = A.EBGaussFieldInitializer.initialize(fields, store, np.zeros(3), "cowan-relativistic-poisson", 1e-3,
                                             (w.cpml, w.cpml, 0), 1e-10, 10000)
    print({k: r[k] for k in REPORT_KEYS}, flush=True)
    assert 1 < r["iterations"] < 10000 and r["relativeResidual"] <= 4e-10
    assert r["totalCharge"] == pytest.approx(-1.602176634e-19 * w.weight * w.particles, rel=1e-12)
    # the box-frame speed of the mean lab momentum under the resonant boost
    from aprl_amd import bunch
    ub = bunch.lab_to_box_u(np.array([[0.0, 0.0, np.sqrt(w.gamma_lab ** 2 - 1.0)]]), w.gamma_boost)
    beta_box = ub[0, 2] / bunch.gamma_from_u(ub)[0]
    print("box beta_z of the mean momentum", beta_box, flush=True)
    assert r["meanBetaZ"] == pytest.approx(beta_box, rel=1e-2)
    ez = fields.component(2)
    assert ez.min() < 0 < ez.max()  # a negative bunch: Ez points inwards at both ends
    assert r["initialElectricEnergy"] > 0 and r["initialMagneticEnergy"] > 0


def test_bending_bunch_rejected_by_the_cowan_positivity_bound():
    import aprl_amd as A
    w = W.BENDING
    fields, window, coupling, store = W.setup_gpu(w, "linear-binomial3", "fp64")
    with pytest.raises(A.AprlError, match="fails the sufficient positive-definiteness bound on axis 2"):
        A.EBGaussFieldInitializer.initialize(fields, store, np.zeros(3), "cowan-relativistic-poisson", 1e-3,
                                             (w.cpml, w.cpml, 0), 1e-10, 10000)
