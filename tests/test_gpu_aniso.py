"""Periodic-box estimators on meshes that are not powers of two and on boxes
whose cells are not cubes, as the reference's own BOSS jobs use them
(RUN_boss.py:85-120: 250 x 460 x 260 on 1350 x 2450 x 1400 Mpc/h and its
siblings), scaled down so the oracle runs in seconds: every FFT length here
goes through the run-time radix plan (radix 5, 13, 23, odd lengths), and the
bispectrum shot noise groups separations by (|i'|, |j'|, |k'|) instead of
|n|^2.  Against the oracle's restatement of calcPowerSpectrumForBOX
(powerspec.hpp:273-332), calcTwoPointCorrelationFunctionForBOX (334-393),
calcBiSpectrumForBOX (bispec.hpp:877-1203) and
calcThreePointCorrelationFunctionForBOX (1205-1424).

Tolerances: mode, shell and point counts bit-exact; P_l within 1e-10 (FP64) /
2e-6 (FP32, measured <= 8e-7) of |P_0| + V/N; xi_l within 1e-10 / 2e-5 of max|xi_0|; B and zeta
within 1e-9 / 5e-6 of their largest value.
"""
import numpy as np
import pytest

import hitomi_amd as hm
from oracle import pyoracle as po

from .conftest import rel_check, report_err

pytestmark = pytest.mark.gpu

# BOSS North zbin 1 divided by 10 (odd n_x), and an isotropic-cell mesh of
# non-power-of-two lengths
MESHES = {
    "boss_scaled": ((25, 46, 26), (135.0, 245.0, 140.0)),
    "boss_scaled_even": ((24, 46, 26), (135.0, 245.0, 140.0)),
    "cubic_cells_odd": ((25, 45, 26), (250.0, 450.0, 260.0)),
    "cubic_cells_mixed": ((36, 40, 30), (360.0, 400.0, 300.0)),
}
SCHEMES = {"NGP": po.NGP, "CIC": po.CIC, "TSC": po.TSC}


def catalogue(L, n, seed, nclump=12, sigma=8.0):
    rng = np.random.default_rng(seed)
    L = np.asarray(L, dtype=np.float64)
    centres = rng.random((nclump, 3)) * L
    k = n // 2
    a = centres[rng.integers(0, nclump, k)] + rng.normal(0.0, sigma, (k, 3))
    b = rng.random((n - k, 3)) * L
    pos = np.mod(np.concatenate([a, b]), L)
    pos[pos >= L] = 0.0
    return np.ascontiguousarray(pos)


@pytest.mark.parametrize("mesh", list(MESHES))
@pytest.mark.parametrize("scheme", list(SCHEMES))
@pytest.mark.parametrize("prec", [hm.FP64, hm.FP32])
def test_aniso_pk(ctx, mesh, scheme, prec):
    n, L = MESHES[mesh]
    N = 20000
    pos = catalogue(L, N, seed=sum(n))
    kw = dict(kmin=0.02, kmax=0.3, n_kbin=15)
    plan = hm.BoxPlan(ctx, n, L, scheme, pk_ells=(0, 2, 4), bk=False, precision=prec, **kw)
    res = plan.execute(pos)
    plan.close()
    V = float(np.prod(L))
    ref0 = None
    for e, l in enumerate((0, 2, 4)):
        p = po.param(n, L, SCHEMES[scheme], ell1=l, ELL=l, **kw)
        ref, nm = po.box_pk(p, pos)
        ref0 = ref if l == 0 else ref0
        assert np.array_equal(res["pk_nmode"], nm.astype(np.int64))
        e_rel = (np.abs(res["pk"][e, :, 0] - ref.real) / (np.abs(ref0.real) + V / N)).max()
        report_err(f"P_{l}", e_rel)
        assert e_rel <= (1e-10 if prec == hm.FP64 else 2e-6)


@pytest.mark.parametrize("mesh", list(MESHES))
@pytest.mark.parametrize("prec", [hm.FP64, hm.FP32])
def test_aniso_xi(ctx, mesh, prec):
    n, L = MESHES[mesh]
    N = 20000
    pos = catalogue(L, N, seed=7 + sum(n))
    rw = dict(rmin=10.0, rmax=60.0, n_rbin=6)
    plan = hm.BoxPlan(ctx, n, L, "TSC", pk_ells=(), bk=False, xi_ells=(0, 2), precision=prec, **rw)
    res = plan.execute(pos)
    plan.close()
    p = po.param(n, L, po.TSC, **rw)
    rb = po.rbin(p)
    d = po.fourier_transform(p, po.box_density_fluctuation(p, pos))
    V = float(np.prod(L))
    ref = []
    for l in (0, 2):
        xi, nm = po.calc_correlation_function(p, d, d, rb, float(N), l, 0)
        ref.append((V / N / N) * (2 * l + 1) * xi.real)
        assert np.array_equal(res["xi_nmode"], nm.astype(np.int64))
    ref = np.array(ref)
    rel_check("xi", res["xi"], ref, 1e-10 if prec == hm.FP64 else 2e-5, np.abs(ref[0]).max())


_BK = {}


def oracle_bk(n, L, pos, kw, ells):
    key = (n, L, ells, pos.shape[0])
    if key not in _BK:
        nk = kw["n_kbin"]
        out = np.zeros((3, nk, nk), complex)
        for ith in range(nk):
            p = po.param(n, L, po.TSC, ith_kbin=ith, ell1=ells[0], ell2=ells[1], ELL=ells[2], **kw)
            out[0, ith], out[1, ith], out[2, ith] = po.box_bk(p, pos)
        _BK[key] = out
    return _BK[key]


@pytest.mark.parametrize("mesh", list(MESHES))
@pytest.mark.parametrize("ells", [(0, 0, 0), (2, 0, 2)])
@pytest.mark.parametrize("mode", ["fp64", "fp64_plain", "fp32"])
def test_aniso_bk(ctx, mesh, ells, mode):
    n, L = MESHES[mesh]
    N = 15000
    pos = catalogue(L, N, seed=11 + sum(n))
    kw = dict(kmin=0.05, kmax=0.2, n_kbin=4)
    prec = hm.FP32 if mode == "fp32" else hm.FP64
    plain = mode == "fp64_plain"
    plan = hm.BoxPlan(ctx, n, L, "TSC", pk_ells=(0,), bk=True, bk_ells=ells, precision=prec, regrid=not plain,
                      prune=not plain, **kw)
    res = plan.execute(pos)
    plan.close()
    ref, ref_bk, ref_sn = oracle_bk(n, L, pos, kw, ells)
    c = lambda a: a[..., 0] + 1j * a[..., 1]
    tol = 1e-9 if prec == hm.FP64 else 5e-6
    report_err("bk_sum", np.abs(c(res["bk_sum"]) - ref_bk).max() / np.abs(ref_bk).max())
    report_err("sn_sum", np.abs(c(res["sn_sum"]) - ref_sn).max() / np.abs(ref_sn).max())
    rel_check("bk", c(res["bk"]), ref, tol)


@pytest.mark.parametrize("mesh", list(MESHES))
@pytest.mark.parametrize("ells", [(0, 0, 0), (1, 1, 0)])
@pytest.mark.parametrize("prec", [hm.FP64, hm.FP32])
def test_aniso_zeta(ctx, mesh, ells, prec):
    n, L = MESHES[mesh]
    N = 8000
    pos = catalogue(L, N, seed=23 + sum(n))
    rw = dict(rmin=20.0, rmax=60.0, n_rbin=3)
    plan = hm.BoxPlan(ctx, n, L, "TSC", pk_ells=(), bk=False, zeta=True, zeta_ells=ells, precision=prec, **rw)
    res = plan.execute(pos)
    plan.close()
    nr = rw["n_rbin"]
    ref = np.zeros((3, nr, nr), complex)
    for ith in range(nr):
        p = po.param(n, L, po.TSC, ith_rbin=ith, ell1=ells[0], ell2=ells[1], ELL=ells[2], **rw)
        ref[0, ith], ref[1, ith], ref[2, ith] = po.box_zeta(p, pos)
    c = lambda a: a[..., 0] + 1j * a[..., 1]
    report_err("zeta_sum", np.abs(c(res["zeta_sum"]) - ref[1]).max() / np.abs(ref[1]).max())
    report_err("zeta_sn", np.abs(c(res["zeta_sn"]) - ref[2]).max() / np.abs(ref[2]).max())
    rel_check("zeta", c(res["zeta"]), ref[0], 1e-9 if prec == hm.FP64 else 5e-6)
