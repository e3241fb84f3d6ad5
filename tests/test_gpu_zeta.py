"""Periodic-box three-point correlation function (hm_box_plan with zeta = 1)
against the oracle's restatement of calcThreePointCorrelationFunctionForBOX
(bispec.hpp:1205-1424, field.hpp:403-445 and 979-1113), every ith_rbin row.

Tolerances (stated per quantity):
  - zeta_save: |d| <= tol * max |zeta_save| over the whole (r1, r2) matrix,
    tol = 1e-9 for FP64 grids and 5e-6 for FP32 grids (measured <= 6e-7; the triple sums run in
    the grid precision within 64-voxel chunks, FP64 across chunks);
  - sn_save (diagonal only, zero elsewhere as in the reference): |d| <= tol *
    max |sn_save|, tol = 1e-10 (FP64) / 2e-6 (FP32, measured <= 3e-7);
  - norm * (zeta - sn): |d| <= tol_zeta * norm * (max|zeta_save| + max|sn_save|).
"""
import numpy as np
import pytest

import hitomi_amd as hm
from oracle import pyoracle as po

from .conftest import clustered_catalogue, rel_check

pytestmark = pytest.mark.gpu

Z_TOL = {hm.FP64: 1e-9, hm.FP32: 5e-6}
SN_TOL = {hm.FP64: 1e-10, hm.FP32: 2e-6}

CASES = [
    # n, L, N, rmin, rmax, n_rbin, (ell1, ell2, ELL)
    (16, 400.0, 3000, 40.0, 120.0, 5, (0, 0, 0)),
    (32, 1000.0, 8000, 30.0, 150.0, 6, (0, 0, 0)),   # shot-noise separations pruned to |r_a| <= 6 cells
    (32, 600.0, 6000, 20.0, 280.0, 5, (1, 1, 0)),    # bins reach half the box: full mesh
    (32, 500.0, 4000, 40.0, 160.0, 4, (2, 0, 2)),
    (16, 400.0, 3000, 40.0, 150.0, 4, (1, 0, 1)),    # odd ell1 + ell2: the odd part of xi in the shot noise
    (16, 400.0, 3000, 40.0, 150.0, 4, (1, 1, 2)),
]

_CACHE = {}


def oracle(case):
    if case not in _CACHE:
        n, L, N, rmin, rmax, nr, (l1, l2, L_) = case
        pos = clustered_catalogue(N, L, seed=n + nr + l1 + 3 * l2, nclump=8, sigma=30.0)
        out = np.zeros((nr, nr), complex)
        zs = np.zeros((nr, nr), complex)
        sn = np.zeros((nr, nr), complex)
        for ith in range(nr):
            p = po.param(n, L, po.TSC, rmin=rmin, rmax=rmax, n_rbin=nr, ith_rbin=ith, ell1=l1, ell2=l2, ELL=L_)
            out[ith], zs[ith], sn[ith] = po.box_zeta(p, pos)
        _CACHE[case] = (pos, out, zs, sn)
    return _CACHE[case]


def cplx(a):
    return a[..., 0] + 1j * a[..., 1]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("mode", ["fp64", "fp64_plain", "fp32"])
def test_zeta_box(ctx, case, mode):
    n, L, N, rmin, rmax, nr, ells = case
    pos, ref, ref_z, ref_sn = oracle(case)
    prec = hm.FP32 if mode == "fp32" else hm.FP64
    plan = hm.BoxPlan(ctx, n, L, "TSC", pk_ells=(), bk=False, n_rbin=nr, rmin=rmin, rmax=rmax, precision=prec,
                      prune=mode != "fp64_plain", zeta=True, zeta_ells=ells)
    res = plan.execute(pos)
    plan.close()
    got_z, got_sn, got = cplx(res["zeta_sum"]), cplx(res["zeta_sn"]), cplx(res["zeta"])
    zscale = np.abs(ref_z).max()
    snscale = np.abs(ref_sn).max()
    assert zscale > 0.0 and snscale > 0.0
    rel_check("zeta_sum", got_z, ref_z, Z_TOL[prec], zscale)
    rel_check("zeta_sn", got_sn, ref_sn, SN_TOL[prec], snscale)
    assert np.all(got_sn[~np.eye(nr, dtype=bool)] == 0.0)
    norm = (L ** 3 / N / N) * (L ** 3 / N)
    rel_check("zeta", got, ref, Z_TOL[prec], norm * (zscale + snscale))


def test_zeta_rejects_forbidden_multipoles(ctx):
    """wigner_3j(1, 0, 0, 0, 0, 0) = 0: the reference refuses it (bispec.hpp:1212-1218)."""
    with pytest.raises(RuntimeError):
        hm.BoxPlan(ctx, 16, 400.0, "TSC", pk_ells=(), bk=False, n_rbin=4, rmin=40.0, rmax=150.0,
                   zeta=True, zeta_ells=(1, 0, 0))
