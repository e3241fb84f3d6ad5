"""Reconstruction (hm_recon_displace, hm_periodic_wrap) and the reconstructed
box estimators (hm_box_plan_execute_recon) against the oracle's restatement
of particle_reconstruction.hpp and powerspec.hpp:396-519.

Tolerances (stated per quantity):
  - displaced positions: |dx| <= tol * max|displacement|, tol = 1e-9 for FP64
    grids and 1e-4 for FP32 grids;
  - the periodic wrap: bit-exact;
  - P_l, xi_l after reconstruction: as the plain box (1e-10 / 1e-5 for P with
    the scale |P_0| + norm * SN, 1e-10 / 2e-5 of max|xi_0| for xi); modes and
    points per bin bit-exact.
"""
import numpy as np
import pytest

import hitomi_amd as hm
from oracle import pyoracle as po

from .conftest import clustered_catalogue, rel_check, uniform_catalogue

pytestmark = pytest.mark.gpu

DISP_TOL = {hm.FP64: 1e-9, hm.FP32: 1e-4}


@pytest.mark.parametrize("prec", [hm.FP64, hm.FP32])
@pytest.mark.parametrize("mode", ["box", "survey"])
def test_recon_displacement(ctx, prec, mode):
    n, L, b1, RG = 32, 1000.0, 2.0, 15.0
    D = clustered_catalogue(20000, L, seed=21, nclump=25, sigma=30.0)
    R = uniform_catalogue(50000, L, seed=22)
    p = po.param(n, L, po.TSC)
    if mode == "box":
        wD = wR = None
        alpha = float(len(D) // len(R))  # Q4: integer division
        norm = L ** 3 / len(D)
    else:
        rng = np.random.default_rng(23)
        wD = 0.5 + rng.random(len(D))
        wR = 0.5 + rng.random(len(R))
        alpha = wD.sum() / wR.sum()
        norm = 0.8 * L ** 3 / wD.sum()
    refD, refR = po.recon_displace(p, D, wD, R, wR, alpha, norm, b1, RG)
    gotD, gotR = ctx.recon_displace(n, L, D, R, alpha, norm, b1, RG, w_d=wD, w_r=wR, precision=prec)
    scale = max(np.abs(refD - D).max(), np.abs(refR - R).max())
    assert scale > 1.0
    assert np.abs(gotD - refD).max() <= DISP_TOL[prec] * scale
    assert np.abs(gotR - refR).max() <= DISP_TOL[prec] * scale


def test_periodic_wrap_bitwise(ctx):
    rng = np.random.default_rng(3)
    L = np.array([1000.0, 800.0, 1200.0])
    pos = rng.uniform(-0.3, 1.3, (5000, 3)) * L
    pos[:3] = [[0.0, 800.0, -0.0], [1000.0, -1e-12, 1200.0 - 1e-9], [-1000.0 * 0.99, 1599.0, 2399.0]]
    got = ctx.periodic_wrap(pos, L)
    ref = pos.copy()
    for a in range(3):
        hi = ref[:, a] >= L[a]
        lo = ref[:, a] < 0.0
        ref[hi, a] -= L[a]
        ref[lo, a] += L[a]
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("prec", [hm.FP64, hm.FP32])
def test_box_recon_pk_xi(ctx, prec):
    n, L, b1, RG = 32, 1000.0, 2.0, 20.0
    D = clustered_catalogue(15000, L, seed=31)
    R = uniform_catalogue(40000, L, seed=32)
    p = po.param(n, L, po.TSC)
    dD, dR = po.recon_displace(p, D, None, R, None, float(len(D) // len(R)), L ** 3 / len(D), b1, RG)
    for a in (dD, dR):
        for ax in range(3):
            a[a[:, ax] >= L, ax] -= L
            a[a[:, ax] < 0.0, ax] += L
    alpha = len(D) / len(R)
    kw = dict(kmin=0.01, kmax=0.15, n_kbin=8)
    rw = dict(rmin=30.0, rmax=200.0, n_rbin=10)
    plan = hm.BoxPlan(ctx, n, L, "TSC", pk_ells=(0, 2), bk=False, xi_ells=(0, 2), precision=prec, **kw, **rw)
    res = plan.execute_recon(dD, dR, alpha)
    plan.close()
    sn = len(D) + alpha ** 2 * len(R)
    norm = L ** 3 / len(D) ** 2
    ref0 = None
    for e, l in enumerate((0, 2)):
        pk = po.param(n, L, po.TSC, ell1=l, ELL=l, **kw)
        ref, nm = po.box_recon_2pt(pk, dD, dR, alpha)
        if l == 0:
            ref0 = ref
        assert np.array_equal(res["pk_nmode"], nm.astype(np.int64))
        scale = np.abs(ref0.real) + norm * sn
        tol = 1e-10 if prec == hm.FP64 else 1e-5
        assert np.all(np.abs(res["pk"][e, :, 0] - ref.real) <= tol * scale)
    xref0 = None
    for e, l in enumerate((0, 2)):
        px = po.param(n, L, po.TSC, ell1=l, ELL=l, **rw)
        ref, nm = po.box_recon_2pt(px, dD, dR, alpha, xi=True)
        if l == 0:
            xref0 = ref
        assert np.array_equal(res["xi_nmode"], nm.astype(np.int64))
        tol = 1e-10 if prec == hm.FP64 else 2e-5
        assert np.abs(res["xi"][e] - ref.real).max() <= tol * np.abs(xref0.real).max()


def _recon_catalogues(n, L, nD, nR, seed, b1=2.0, RG=20.0):
    D = clustered_catalogue(nD, L, seed=seed)
    R = uniform_catalogue(nR, L, seed=seed + 1)
    p = po.param(n, L, po.TSC)
    dD, dR = po.recon_displace(p, D, None, R, None, float(len(D) // len(R)), L ** 3 / len(D), b1, RG)
    for a in (dD, dR):
        for ax in range(3):
            a[a[:, ax] >= L, ax] -= L
            a[a[:, ax] < 0.0, ax] += L
    return dD, dR, len(D) / len(R)


@pytest.mark.parametrize("prec", [hm.FP64, hm.FP32])
@pytest.mark.parametrize("ells", [(0, 0, 0), (1, 1, 0), (2, 0, 2)])
def test_box_recon_bispectrum(ctx, prec, ells):
    """bispec.hpp:1426-1753 for every ith_kbin: |dB| <= tol * max|B| (1e-9 FP64, 5e-6 FP32), the bk and sn
    sums separately with the same bound; shell mode counts bit-exact."""
    n, L = 32, 1000.0
    dD, dR, alpha = _recon_catalogues(n, L, 12000, 30000, seed=61)
    kw = dict(kmin=0.02, kmax=0.1, n_kbin=5)
    plan = hm.BoxPlan(ctx, n, L, "TSC", pk_ells=(0,), bk=True, bk_ells=ells, precision=prec, **kw)
    res = plan.execute_recon(dD, dR, alpha)
    plan.close()
    nk = kw["n_kbin"]
    ref = np.zeros((nk, nk), complex)
    ref_bk = np.zeros((nk, nk), complex)
    ref_sn = np.zeros((nk, nk), complex)
    for ith in range(nk):
        p = po.param(n, L, po.TSC, ell1=ells[0], ell2=ells[1], ELL=ells[2], ith_kbin=ith, **kw)
        ref[ith], ref_bk[ith], ref_sn[ith] = po.box_recon_bk(p, dD, dR, alpha)
    tol = 1e-9 if prec == hm.FP64 else 5e-6
    c = lambda a: a[..., 0] + 1j * a[..., 1]
    for name, got, want in (("bk", c(res["bk"]), ref), ("bk_sum", c(res["bk_sum"]), ref_bk),
                            ("sn_sum", c(res["sn_sum"]), ref_sn)):
        rel_check(name, got, want, tol)


@pytest.mark.parametrize("prec", [hm.FP64, hm.FP32])
@pytest.mark.parametrize("ells", [(0, 0, 0), (1, 1, 0)])
def test_box_recon_zeta(ctx, prec, ells):
    """bispec.hpp:1755-1975 for every ith_rbin: zeta_save within 1e-9 (FP64) / 1e-4 (FP32) of its maximum,
    the shot noise within 1e-10 / 2e-5 of its maximum."""
    n, L = 16, 500.0
    dD, dR, alpha = _recon_catalogues(n, L, 4000, 10000, seed=67, RG=25.0)
    rw = dict(rmin=40.0, rmax=150.0, n_rbin=4)
    plan = hm.BoxPlan(ctx, n, L, "TSC", pk_ells=(), bk=False, zeta=True, zeta_ells=ells, precision=prec, **rw)
    res = plan.execute_recon(dD, dR, alpha)
    plan.close()
    nr = rw["n_rbin"]
    ref_z = np.zeros((nr, nr), complex)
    ref_sn = np.zeros((nr, nr), complex)
    for ith in range(nr):
        p = po.param(n, L, po.TSC, ell1=ells[0], ell2=ells[1], ELL=ells[2], ith_rbin=ith, **rw)
        _, ref_z[ith], ref_sn[ith] = po.box_recon_zeta(p, dD, dR, alpha)
    c = lambda a: a[..., 0] + 1j * a[..., 1]
    zt, st = (1e-9, 1e-10) if prec == hm.FP64 else (1e-4, 2e-5)
    assert np.abs(c(res["zeta_sum"]) - ref_z).max() <= zt * np.abs(ref_z).max()
    assert np.abs(c(res["zeta_sn"]) - ref_sn).max() <= st * np.abs(ref_sn).max()
