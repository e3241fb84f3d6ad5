"""Survey-geometry mocks and the survey estimators at the reference's sizes.

  - hm_mock_sky_randoms against the oracle's restatement (same Philox
    counters): positions within 1e-12 of rmax (FP64 cbrt/sin/cos);
  - the survey P_l workload at full size: 512^3 TSC, L = 2000 Mpc/h,
    10^6 Zel'dovich data in the shell 1200-1800 Mpc/h times RA 0-60 deg,
    Dec -30..30 deg, 5 x 10^7 randoms in the same patch, FKP weights
    1/(1 + nbar 10^4): mode counts bit-exact against the oracle's count-only
    restatement of field.hpp:793-837; FP32 P_0, P_2, P_4 within 1e-5 of
    |P_0| of the FP64 plan (whose arithmetic the survey parity tests pin
    against the oracle at smaller meshes);
  - every mesh / box pair of the reference's BOSS jobs (RUN_boss.py:85-120,
    non-power-of-two, anisotropic cells): mode counts bit-exact, FP32 against
    FP64 as above, for P_0 and P_2; B_000 on the smallest pair, FP32 against
    FP64 within 5e-6 of max|B|.
"""
import numpy as np
import pytest

import hitomi_amd as hm
from hitomi_amd import mocks
from oracle import pyoracle as po

from .conftest import report_err

pytestmark = pytest.mark.gpu

C2_PATCH = dict(rmin=1200.0, rmax=1800.0, ra_min=0.0, ra_max=60.0, dec_min=-30.0, dec_max=30.0)
# RUN_boss.py:85-120 (North zbin 1, 3; South zbin 1, 3)
BOSS = [((250, 460, 260), (1350.0, 2450.0, 1400.0)), ((340, 650, 360), (1800.0, 3400.0, 1900.0)),
        ((190, 360, 210), (1000.0, 1900.0, 1100.0)), ((190, 500, 280), (1000.0, 2600.0, 1500.0))]
BOSS_PATCH = dict(rmin=500.0, rmax=800.0, ra_min=0.0, ra_max=40.0, dec_min=-20.0, dec_max=20.0)


def test_sky_randoms_match_oracle(ctx):
    patch = hm.sky_patch(**C2_PATCH)
    n = 200_000
    d = hm.mock_sky_randoms(ctx, patch, 12345, n)
    ctx.sync()
    got = d.download()
    d.free()
    ref = po.sky_randoms([C2_PATCH[k] for k in ("rmin", "rmax", "ra_min", "ra_max", "dec_min", "dec_max")], 12345, n)
    assert np.abs(got - ref).max() <= 1e-12 * C2_PATCH["rmax"]
    assert mocks.in_sky_patch(got, patch).all()


@pytest.fixture(scope="module")
def c2(ctx):
    patch = hm.sky_patch(**C2_PATCH)
    cat = mocks.survey_catalogue(ctx, patch, 1_000_000, 50_000_000, seed=2024)
    yield cat
    cat["posR"].free()
    cat["wR"].free()


def test_c2_catalogue_shape(c2):
    patch = hm.sky_patch(**C2_PATCH)
    assert c2["posD"].shape == (1_000_000, 3) and c2["posR"].shape == (50_000_000, 3)
    assert mocks.in_sky_patch(c2["posD"], patch).all()
    w = 1.0 / (1.0 + c2["nbar"] * 1e4)
    assert np.all(c2["wD"] == w)
    assert c2["nbar"] == pytest.approx(1e6 / mocks.sky_patch_volume(patch))


def test_c2_survey_pk_full_size(ctx, c2):
    n, L = 512, 2000.0
    kw = dict(kmin=0.015, kmax=0.205, n_kbin=20)
    ref_nm, _ = po.mode_counts(po.param(n, L, po.TSC, **kw))
    out = {}
    for ell in (0, 2, 4):
        for prec in (hm.FP32, hm.FP64):
            plan = hm.SurveyPlan(ctx, n, L, "TSC", ell=ell, precision=prec, **kw)
            pk, nm, info = plan.execute(c2["posD"], c2["wD"], c2["posR"], c2["wR"])
            plan.close()
            assert np.array_equal(nm, ref_nm), (ell, prec)
            out[(ell, prec)] = pk
    p0 = np.abs(out[(0, hm.FP64)].real)
    assert np.all(p0 > 0)
    for ell in (0, 2, 4):
        e = (np.abs(out[(ell, hm.FP32)] - out[(ell, hm.FP64)]) / p0).max()
        report_err(f"C2 P_{ell} fp32 vs fp64", e)
        assert e <= 1e-5


@pytest.mark.parametrize("mesh", BOSS, ids=["NGC_z1", "NGC_z3", "SGC_z1", "SGC_z3"])
def test_boss_meshes_survey_pk(ctx, mesh):
    n, L = mesh
    patch = hm.sky_patch(**BOSS_PATCH)
    cat = mocks.survey_catalogue(ctx, patch, 100_000, 1_000_000, seed=sum(n), n_side=128, randoms_on_device=False)
    kw = dict(kmin=0.01, kmax=0.2, n_kbin=20)
    ref_nm, _ = po.mode_counts(po.param(n, L, po.TSC, **kw))
    out = {}
    for ell in (0, 2):
        for prec in (hm.FP32, hm.FP64):
            plan = hm.SurveyPlan(ctx, n, L, "TSC", ell=ell, precision=prec, **kw)
            pk, nm, _ = plan.execute(cat["posD"], cat["wD"], cat["posR"], cat["wR"])
            plan.close()
            assert np.array_equal(nm, ref_nm), (ell, prec)
            out[(ell, prec)] = pk
    p0 = np.abs(out[(0, hm.FP64)].real)
    for ell in (0, 2):
        e = (np.abs(out[(ell, hm.FP32)] - out[(ell, hm.FP64)]) / p0).max()
        report_err(f"BOSS {n} P_{ell} fp32 vs fp64", e)
        assert e <= 1e-5


def test_boss_mesh_survey_bk(ctx):
    n, L = BOSS[2]
    patch = hm.sky_patch(**BOSS_PATCH)
    cat = mocks.survey_catalogue(ctx, patch, 100_000, 1_000_000, seed=7, n_side=128, randoms_on_device=False)
    kw = dict(kmin=0.02, kmax=0.1, n_kbin=5)
    out = {}
    for prec in (hm.FP32, hm.FP64):
        plan = hm.SurveyPlan(ctx, n, L, "TSC", measure="bk", ell=0, precision=prec, **kw)
        bk, nm, _ = plan.execute(cat["posD"], cat["wD"], cat["posR"], cat["wR"])
        plan.close()
        out[prec] = bk
    e = np.abs(out[hm.FP32] - out[hm.FP64]).max() / np.abs(out[hm.FP64]).max()
    report_err("BOSS SGC_z1 B_000 fp32 vs fp64", e)
    assert e <= 5e-6
