"""Periodic-box P_l and B_000 (hm_box_plan) against the oracle's restatement of
calcPowerSpectrumForBOX (powerspec.hpp:273-332) and calcBiSpectrumForBOX
(bispec.hpp:877-1203, one oracle job per ith_kbin).

Tolerances (stated per quantity):
  - mode counts (P bins and B shells): bit-exact;
  - P_l: |dP_l(k)| <= tol * (|P_0(k)| + V/N), tol = 1e-10 for FP64 grids and
    1e-6 for FP32 grids (FP64 bin sums in both; measured <= 1.5e-7);
  - B: |dB(k1,k2)| <= tol * max|B| with tol = 1e-9 (FP64) / 5e-6 (FP32, measured
    <= 2.7e-6 for the odd multipole (1, 0, 1), <= 4e-7 otherwise); the
    bk and sn sums are checked separately with the same tolerances.
"""
import numpy as np
import pytest

import hitomi_amd as hm
from oracle import pyoracle as po

from .conftest import clustered_catalogue, rel_check, report_err, uniform_catalogue

pytestmark = pytest.mark.gpu

PK_TOL = {hm.FP64: 1e-10, hm.FP32: 1e-6}
BK_TOL = {hm.FP64: 1e-9, hm.FP32: 5e-6}


def oracle_pk(n, L, pos, kmin, kmax, nk, ells=(0, 2, 4)):
    out = []
    for l in ells:
        p = po.param(n, L, po.TSC, kmin=kmin, kmax=kmax, n_kbin=nk, ell1=l, ell2=0, ELL=l)
        pk, nm = po.box_pk(p, pos)
        out.append(pk)
    return np.array(out), nm


def check_pk(res, ref_pk, ref_nm, prec, V, N):
    assert np.array_equal(res["pk_nmode"], ref_nm.astype(np.int64))
    got = res["pk"][..., 0] + 1j * res["pk"][..., 1]
    scale = np.abs(ref_pk[0].real) + V / N
    e = (np.abs(got - ref_pk) / scale[None, :]).max()
    report_err("pk", e)
    assert e <= PK_TOL[prec]


@pytest.mark.parametrize("prec", [hm.FP64, hm.FP32])
def test_pk_config1_catalogue(ctx, prec):
    """128^3, 2e5 particles, L = 1000, z-RSD sigma = 5, bins 0.005..0.195 (20)."""
    n, L, N = 128, 1000.0, 200000
    pos = uniform_catalogue(N, L, seed=1, rsd_sigma=5.0)
    ref_pk, ref_nm = oracle_pk(n, L, pos, 0.005, 0.195, 20)
    plan = hm.BoxPlan(ctx, n, L, "TSC", n_kbin=20, kmin=0.005, kmax=0.195, pk_ells=(0, 2, 4), bk=False, precision=prec)
    res = plan.execute(pos)
    check_pk(res, ref_pk, ref_nm, prec, L ** 3, N)


@pytest.mark.parametrize("prec", [hm.FP64, hm.FP32])
def test_pk_config1_catalogue_interlaced(ctx, prec):
    """The same catalogue and bins with TSC interlacing (no counterpart in the
    reference) against the oracle's restatement of it: mode counts bit-exact;
    |dP_l| <= 1e-10 (FP64) / 1e-5 (FP32) of |P_0| + V/N."""
    n, L, N = 128, 1000.0, 200000
    pos = uniform_catalogue(N, L, seed=1, rsd_sigma=5.0)
    kw = dict(n_kbin=20, kmin=0.005, kmax=0.195)
    plan = hm.BoxPlan(ctx, n, L, "TSC", pk_ells=(0, 2, 4), bk=False, precision=prec, interlace=True, **kw)
    res = plan.execute(pos)
    plan.close()
    tol = 1e-10 if prec == hm.FP64 else 1e-5
    ref0 = None
    for e, l in enumerate((0, 2, 4)):
        ref, nm = po.box_interlaced_2pt(po.param(n, L, po.TSC, ell1=l, ELL=l, **kw), pos)
        ref0 = ref if l == 0 else ref0
        assert np.array_equal(res["pk_nmode"], nm.astype(np.int64))
        err = (np.abs(res["pk"][e, :, 0] - ref.real) / (np.abs(ref0.real) + L ** 3 / N)).max()
        report_err(f"interlaced 128^3 P_{l}", err)
        assert err <= tol


@pytest.mark.parametrize("prec", [hm.FP64, hm.FP32])
def test_pk_low_k_pruned(ctx, prec):
    """Bins far below Nyquist: the binning visits only the low-k cube."""
    n, L, N = 64, 2000.0, 30000
    pos = clustered_catalogue(N, L, seed=2)
    ref_pk, ref_nm = oracle_pk(n, L, pos, 0.01, 0.04, 7)
    plan = hm.BoxPlan(ctx, n, L, "TSC", n_kbin=7, kmin=0.01, kmax=0.04, pk_ells=(0, 2, 4), bk=False, precision=prec)
    check_pk(plan.execute(pos), ref_pk, ref_nm, prec, L ** 3, N)


_ORACLE_CACHE = {}


def oracle_bk_all(n, L, pos, kmin, kmax, nk):
    key = (n, L, pos.shape[0], kmin, kmax, nk, float(pos.sum()))
    if key not in _ORACLE_CACHE:
        _ORACLE_CACHE[key] = _oracle_bk_all(n, L, pos, kmin, kmax, nk)
    return _ORACLE_CACHE[key]


def _oracle_bk_all(n, L, pos, kmin, kmax, nk):
    out = np.zeros((nk, nk), complex)
    bks = np.zeros((nk, nk), complex)
    sns = np.zeros((nk, nk), complex)
    for ith in range(nk):
        p = po.param(n, L, po.TSC, kmin=kmin, kmax=kmax, n_kbin=nk, ith_kbin=ith)
        o, b, s = po.box_bk(p, pos)
        out[ith], bks[ith], sns[ith] = o, b, s
    return out, bks, sns


def shell_counts(n, L, kmin, kmax, nk):
    p = po.param(n, L, po.TSC, kmin=kmin, kmax=kmax, n_kbin=nk)
    kb = po.kbin(p)
    one = np.ones((n, n, n), complex)
    return np.array([po.ifft_for_bispectrum(p, one, k, kb[1] - kb[0])[1] for k in kb])


BK_CASES = [
    # n, L, N, kmin, kmax, nk, expected triple mesh
    (64, 2000.0, 20000, 0.01, 0.025, 4, 40),   # exact re-grid 64 -> 40 (4 s + 1 = 37)
    (32, 1000.0, 20000, 0.02, 0.08, 4, 32),    # shells pruned, no re-grid
    (64, 2000.0, 20000, 0.005, 0.05, 10, 64),  # 10 shells: three 4-shell groups in the fused triple kernel
    (64, 1500.0, 15000, 0.01, 0.04, 7, 48),    # 7 shells: a partial group; re-grid 64 -> 48
]


@pytest.mark.parametrize("case", BK_CASES)
@pytest.mark.parametrize("mode", ["fp64", "fp64_plain", "fp32"])
def test_bk000_all_pairs(ctx, case, mode):
    n, L, N, kmin, kmax, nk, M = case
    pos = clustered_catalogue(N, L, seed=n + nk, nclump=30, sigma=30.0)
    ref, ref_bk, ref_sn = oracle_bk_all(n, L, pos, kmin, kmax, nk)
    prec = hm.FP32 if mode == "fp32" else hm.FP64
    plain = mode == "fp64_plain"
    plan = hm.BoxPlan(ctx, n, L, "TSC", n_kbin=nk, kmin=kmin, kmax=kmax, pk_ells=(0,), bk=True, precision=prec,
                      regrid=not plain, prune=not plain)
    assert plan.triple_mesh == ((n, n, n) if plain else (M, M, M))
    res = plan.execute(pos)
    assert np.array_equal(res["shell_nmode"], shell_counts(n, L, kmin, kmax, nk))
    c = lambda a: a[..., 0] + 1j * a[..., 1]
    tol = BK_TOL[prec]
    rel_check("bk_sum", c(res["bk_sum"]), ref_bk, tol)
    rel_check("sn_sum", c(res["sn_sum"]), ref_sn, tol)
    rel_check("bk", c(res["bk"]), ref, tol)


def test_fp32_pipeline_bitwise_deterministic(ctx):
    n, L, N = 64, 2000.0, 40000
    pos = clustered_catalogue(N, L, seed=5)
    plan = hm.BoxPlan(ctx, n, L, "TSC", n_kbin=4, kmin=0.01, kmax=0.025, pk_ells=(0, 2, 4), bk=True, precision=hm.FP32)
    a = plan.execute(pos)
    perm = np.random.default_rng(0).permutation(N)
    b = plan.execute(pos[perm])
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_metric_mesh_512_properties(ctx):
    """512^3, L = 2000, 4e6 particles, 20 bins 0.015..0.205: FP32 P_l agree with the
    FP64 plan to 1e-6 (P_0 + SN); B within 5e-6 of max|B|; B is symmetric in (k1, k2); repeat runs
    are bitwise equal."""
    n, L, N = 512, 2000.0, 4_000_000
    pos = clustered_catalogue(N, L, seed=42, nclump=400, sigma=15.0)
    kw = dict(n_kbin=20, kmin=0.015, kmax=0.205, pk_ells=(0, 2, 4), bk=True)
    p32 = hm.BoxPlan(ctx, n, L, "TSC", precision=hm.FP32, **kw)
    dpos = ctx.array(pos)
    r1 = p32.execute(dpos)
    r2 = p32.execute(dpos)
    for k in r1:
        assert np.array_equal(r1[k], r2[k]), k
    p32.close()
    p64 = hm.BoxPlan(ctx, n, L, "TSC", precision=hm.FP64, **kw)
    r64 = p64.execute(dpos)
    p64.close()
    assert np.array_equal(r1["pk_nmode"], r64["pk_nmode"])
    scale = np.abs(r64["pk"][0, :, 0]) + L ** 3 / N
    assert np.all(np.abs(r1["pk"][..., 0] - r64["pk"][..., 0]) <= 1e-6 * scale[None, :])
    b = r1["bk"][..., 0]
    assert np.allclose(b, b.T, rtol=0, atol=0)
    b64 = r64["bk"][..., 0]
    rel_check("B_000 fp32 vs fp64", b, b64, 5e-6)


_GEN_CACHE = {}


def oracle_bk_general(n, L, pos, kmin, kmax, nk, ells):
    key = (n, L, pos.shape[0], kmin, kmax, nk, ells, float(pos.sum()))
    if key not in _GEN_CACHE:
        out = np.zeros((nk, nk), complex)
        bks = np.zeros((nk, nk), complex)
        sns = np.zeros((nk, nk), complex)
        for ith in range(nk):
            p = po.param(n, L, po.TSC, kmin=kmin, kmax=kmax, n_kbin=nk, ith_kbin=ith, ell1=ells[0], ell2=ells[1],
                         ELL=ells[2])
            out[ith], bks[ith], sns[ith] = po.box_bk(p, pos)
        _GEN_CACHE[key] = (out, bks, sns)
    return _GEN_CACHE[key]


GEN_CASES = [
    # n, L, N, kmin, kmax, nk, ells, expected triple mesh
    (32, 1000.0, 20000, 0.02, 0.08, 4, (1, 1, 0), 32),
    (32, 1000.0, 20000, 0.02, 0.08, 4, (2, 0, 2), 32),
    (32, 1000.0, 20000, 0.02, 0.08, 4, (0, 2, 2), 32),
    (32, 1000.0, 20000, 0.02, 0.08, 4, (1, 0, 1), 32),   # odd l1 + l2: imaginary shot noise
    (64, 2000.0, 20000, 0.01, 0.025, 4, (1, 1, 2), 40),  # re-grid 64 -> 40
]


@pytest.mark.parametrize("case", GEN_CASES)
@pytest.mark.parametrize("mode", ["fp64", "fp64_plain", "fp32"])
def test_bk_general_multipole(ctx, case, mode):
    """B_{l1 l2 L}(k1, k2) for every ordered pair: bk and sn sums and the normalised
    output against the oracle (tolerances as for B_000)."""
    n, L, N, kmin, kmax, nk, ells, M = case
    pos = clustered_catalogue(N, L, seed=n + nk + sum(ells), nclump=30, sigma=30.0)
    ref, ref_bk, ref_sn = oracle_bk_general(n, L, pos, kmin, kmax, nk, ells)
    prec = hm.FP32 if mode == "fp32" else hm.FP64
    plain = mode == "fp64_plain"
    plan = hm.BoxPlan(ctx, n, L, "TSC", n_kbin=nk, kmin=kmin, kmax=kmax, pk_ells=(0,), bk=True, precision=prec,
                      regrid=not plain, prune=not plain, bk_ells=ells)
    assert plan.triple_mesh == ((n, n, n) if plain else (M, M, M))
    res = plan.execute(pos)
    c = lambda a: a[..., 0] + 1j * a[..., 1]
    tol = BK_TOL[prec]
    rel_check("bk_sum", c(res["bk_sum"]), ref_bk, tol)
    rel_check("sn_sum", c(res["sn_sum"]), ref_sn, tol)
    rel_check("bk", c(res["bk"]), ref, tol)


@pytest.mark.parametrize("prec", [hm.FP64, hm.FP32])
def test_interlaced_pk_xi(ctx, prec):
    """Interlacing (no counterpart in the reference): P_l and xi_l of (d_a + exp(i k.h) d_b)/2 with the shot
    noise N W^2 against the oracle's restatement; tolerances (P: 1e-10 / 1e-5 of
    |P_0| + V/N, xi: 1e-10 / 2e-5 of max|xi_0|), mode and point counts bit-exact."""
    n, L, N = 32, 1000.0, 30000
    pos = clustered_catalogue(N, L, seed=91, nclump=25, sigma=25.0)
    kw = dict(kmin=0.02, kmax=0.3, n_kbin=14)
    rw = dict(rmin=40.0, rmax=200.0, n_rbin=9)
    plan = hm.BoxPlan(ctx, n, L, "TSC", pk_ells=(0, 2), bk=False, xi_ells=(0, 2), precision=prec, interlace=True,
                      **kw, **rw)
    res = plan.execute(pos)
    plan.close()
    tol = 1e-10 if prec == hm.FP64 else 1e-5
    ref0 = None
    for e, l in enumerate((0, 2)):
        p = po.param(n, L, po.TSC, ell1=l, ELL=l, **kw)
        ref, nm = po.box_interlaced_2pt(p, pos)
        ref0 = ref if l == 0 else ref0
        assert np.array_equal(res["pk_nmode"], nm.astype(np.int64))
        e_rel = (np.abs(res["pk"][e, :, 0] - ref.real) / (np.abs(ref0.real) + L ** 3 / N)).max()
        report_err(f"interlaced P_{l}", e_rel)
        assert e_rel <= tol
    xtol = 1e-10 if prec == hm.FP64 else 2e-5
    x0 = None
    for e, l in enumerate((0, 2)):
        p = po.param(n, L, po.TSC, ell1=l, ELL=l, **rw)
        ref, nm = po.box_interlaced_2pt(p, pos, xi=True)
        x0 = ref if l == 0 else x0
        assert np.array_equal(res["xi_nmode"], nm.astype(np.int64))
        assert np.abs(res["xi"][e] - ref.real).max() <= xtol * np.abs(x0.real).max()


def test_interlace_rejected_for_bispectrum(ctx):
    with pytest.raises(RuntimeError):
        hm.BoxPlan(ctx, 32, 1000.0, "TSC", bk=True, interlace=True, kmin=0.02, kmax=0.1, n_kbin=5)
