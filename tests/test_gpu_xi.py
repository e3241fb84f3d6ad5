"""Periodic-box xi_l (hm_box_plan with xi_ells) against the oracle's
restatement of calcTwoPointCorrelationFunctionForBOX (powerspec.hpp:334-393)
and calcCorrelationFunction (field.hpp:854-977).

Tolerances (stated per quantity):
  - points per r-bin: bit-exact;
  - xi_l: |dxi_l(r)| <= tol * max_r |xi_0(r)|, tol = 1e-10 for FP64 grids and
    2e-5 for FP32 grids (FP64 bin sums in both).  Odd l are pure rounding
    noise in the reference (the even field against an odd Legendre
    polynomial over a symmetric shell) and are held to the same bound.
"""
import numpy as np
import pytest

import hitomi_amd as hm
from oracle import pyoracle as po

from .conftest import clustered_catalogue

pytestmark = pytest.mark.gpu

XI_TOL = {hm.FP64: 1e-10, hm.FP32: 2e-5}


def oracle_xi(n, L, pos, rmin, rmax, nr, ells):
    p = po.param(n, L, po.TSC, rmin=rmin, rmax=rmax, n_rbin=nr)
    rb = po.rbin(p)
    d = po.fourier_transform(p, po.box_density_fluctuation(p, pos))
    N = pos.shape[0]
    out, nm = [], None
    for l in ells:
        xi, nm = po.calc_correlation_function(p, d, d, rb, float(N), l, 0)
        out.append((L ** 3 / N / N) * (2 * l + 1) * xi.real)
    return np.array(out), nm


XI_CASES = [
    # n, L, N, rmin, rmax, n_rbin, ells
    (64, 1000.0, 30000, 20.0, 150.0, 14, (0, 2, 4)),  # separations pruned to |r_a| <= 151 Mpc/h
    (32, 600.0, 20000, 10.0, 290.0, 15, (0, 2)),      # bins reach the box half-diagonal: full mesh
    (64, 2000.0, 20000, 50.0, 400.0, 8, (0, 1, 3)),   # odd multipoles
]


@pytest.mark.parametrize("case", XI_CASES)
@pytest.mark.parametrize("mode", ["fp64", "fp64_plain", "fp32"])
def test_xi_box(ctx, case, mode):
    n, L, N, rmin, rmax, nr, ells = case
    pos = clustered_catalogue(N, L, seed=n + nr, nclump=25, sigma=25.0)
    ref, ref_nm = oracle_xi(n, L, pos, rmin, rmax, nr, ells)
    prec = hm.FP32 if mode == "fp32" else hm.FP64
    plan = hm.BoxPlan(ctx, n, L, "TSC", pk_ells=(), bk=False, xi_ells=ells, n_rbin=nr, rmin=rmin, rmax=rmax,
                      precision=prec, prune=mode != "fp64_plain")
    res = plan.execute(pos)
    assert np.array_equal(res["xi_nmode"], ref_nm.astype(np.int64))
    scale = np.abs(ref[0]).max()
    err = np.abs(res["xi"] - ref).max()
    assert err <= XI_TOL[prec] * scale, (err, scale)


def test_pk_xi_bk_in_one_plan(ctx):
    """P_l, xi_l and B_000 from one plan: the xi transform shares its buffer
    with the bispectrum shot-noise transform and must not disturb it."""
    n, L, N, nk = 32, 1000.0, 20000, 4
    pos = clustered_catalogue(N, L, seed=11)
    plan = hm.BoxPlan(ctx, n, L, "TSC", n_kbin=nk, kmin=0.02, kmax=0.08, pk_ells=(0, 2), bk=True, xi_ells=(0, 2),
                      n_rbin=10, rmin=40.0, rmax=220.0, precision=hm.FP64)
    res = plan.execute(pos)
    ref_xi, ref_nm = oracle_xi(n, L, pos, 40.0, 220.0, 10, (0, 2))
    assert np.array_equal(res["xi_nmode"], ref_nm.astype(np.int64))
    assert np.abs(res["xi"] - ref_xi).max() <= 1e-10 * np.abs(ref_xi[0]).max()
    for ith in range(nk):
        p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.08, n_kbin=nk, ith_kbin=ith)
        out, _, _ = po.box_bk(p, pos)
        got = res["bk"][ith, :, 0] + 1j * res["bk"][ith, :, 1]
        assert np.abs(got - out).max() <= 1e-9 * np.abs(out).max()
    for e, l in enumerate((0, 2)):
        p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.08, n_kbin=nk, ell1=l, ELL=l)
        pk, nm = po.box_pk(p, pos)
        got = res["pk"][e, :, 0]
        assert np.abs(got - pk.real).max() <= 1e-10 * (np.abs(pk.real).max() + L ** 3 / N)
