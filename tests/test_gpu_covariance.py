"""Covariance stage on the device (hm_cov, hitomi_amd.covariance) against
numpy, as calc_data_and_cov.py computes it:

  - synthetic vectors of the reference's layout (xi_0,2,4 on 25 bins, four
    zeta multipoles on 13 x 13 bins), pushed from the host, as block dicts
    and from device memory: the stored vectors equal numpy's v / w bit for
    bit; the mean and the N - 1 covariance equal numpy.cov of the normalised
    vectors within 1e-12 of the largest variance; each block pair equals
    np.cov(X_a, X_b)[0:n_a, n_a:n_a + n_b] (:288-307);
  - end to end on survey-geometry mocks at 64^3: xi_l and zeta_l1l2L of each
    mock through the survey plans, the window monopoles from window2PCF /
    window3PCF on the randoms, the stage against numpy on the same values;
    the same through the files in the drop-in programs' formats
    (from_measurements), against numpy on the values the files hold.
"""
import os

import numpy as np
import pytest

import hitomi_amd as hm
from hitomi_amd import covariance as cv
from hitomi_amd import mocks

pytestmark = pytest.mark.gpu


def np_blocks(X, names, sizes):
    off = np.concatenate([[0], np.cumsum(sizes)])
    return {n: X[:, off[i]:off[i + 1]] for i, n in enumerate(names)}


def check_against_numpy(c, X):
    mean, cov = c.result()
    ref = np.cov(X.T)
    scale = np.diag(ref).max()
    assert np.abs(cov - ref).max() <= 1e-12 * scale
    assert np.abs(mean - X.mean(0)).max() <= 1e-13 * np.abs(X).max()
    blk = np_blocks(X, c.names, c.sizes)
    for a in c.names:
        for b in c.names:
            na, nb = blk[a].shape[1], blk[b].shape[1]
            want = np.cov(blk[a].T, blk[b].T)[0:na, na:na + nb]  # calc_data_and_cov.py:288
            assert np.abs(c.block(a, b) - want).max() <= 1e-12 * scale, (a, b)
        assert np.allclose(c.std(a), np.std(blk[a], axis=0, ddof=1), rtol=1e-10, atol=0)


def test_synthetic_vectors_reference_layout(ctx):
    rng = np.random.default_rng(11)
    blocks = cv.blocks_for(25, 13)
    w0 = 1.0 + rng.random(25)
    w000 = 0.5 + rng.random((13, 13))
    c = cv.Covariance(ctx, blocks, cv.windows_for(w0, w000), keep_rows=64)
    n = 40
    raw = rng.standard_normal((n, c.dim)) * np.linspace(1.0, 50.0, c.dim) + 100.0
    w = c.norm
    for i, v in enumerate(raw):
        if i % 3 == 0:
            c.push(v)
        elif i % 3 == 1:
            parts = {nm: v[c.slices[nm]] for nm in c.names}
            parts["zeta000"] = parts["zeta000"].reshape(13, 13)
            c.push(parts)
        else:
            d = ctx.array(v.copy())
            c.push(d)
            ctx.sync()
            d.free()
    assert c.count == n
    X = raw / w
    assert np.array_equal(c.vectors(), X)
    check_against_numpy(c, X)
    c.reset()
    for v in raw[:5]:
        c.push(v)
    assert np.abs(c.result()[1] - np.cov(X[:5].T)).max() <= 1e-12 * np.diag(np.cov(X[:5].T)).max()
    c.close()


def test_rejects_bad_input(ctx):
    with pytest.raises(hm.HitomiError):
        cv.Covariance(ctx, [("xi0", 3)], {"xi0": np.array([1.0, 0.0, 2.0])})
    c = cv.Covariance(ctx, [("xi0", 3)])
    c.push(np.ones(3))
    with pytest.raises(hm.HitomiError):
        c.result()
    with pytest.raises(ValueError):
        c.push(np.ones(4))
    c.close()


PATCH = dict(rmin=200.0, rmax=400.0, ra_min=0.0, ra_max=40.0, dec_min=-20.0, dec_max=20.0)
XI_KW = dict(n_rbin=10, rmin=20.0, rmax=110.0)
Z_KW = dict(n_rbin=5, rmin=20.0, rmax=100.0)


def measure_mock(ctx, cat, n, L):
    out = {}
    for l in cv.XI_ELLS:
        p = hm.SurveyPlan(ctx, n, L, "TSC", measure="2PCF", ell=l, **XI_KW)
        xi, _, _ = p.execute(cat["posD"], cat["wD"], cat["posR"], cat["wR"])
        p.close()
        out[cv.xi_name(l)] = xi.real
    for (l1, l2, L_) in cv.ZETA_MULTIPOLES:
        p = hm.SurveyPlan(ctx, n, L, "TSC", measure="3PCF", ell=L_, ell1=l1, ell2=l2, **Z_KW)
        z, _, _ = p.execute(cat["posD"], cat["wD"], cat["posR"], cat["wR"])
        p.close()
        out[cv.zeta_name(l1, l2, L_)] = z.real
    return out


def measure_windows(ctx, cat, n, L):
    p = hm.SurveyPlan(ctx, n, L, "TSC", measure="window2PCF", ell=0, **XI_KW)
    w0, _, _ = p.execute(cat["posD"], cat["wD"], cat["posR"], cat["wR"])
    p.close()
    p = hm.SurveyPlan(ctx, n, L, "TSC", measure="window3PCF", ell=0, ell1=0, ell2=0, **Z_KW)
    w000, _, _ = p.execute(cat["posD"], cat["wD"], cat["posR"], cat["wR"])
    p.close()
    return w0.real, w000.real


def test_survey_mocks_end_to_end(ctx, tmp_path):
    n, L = 64, 600.0
    patch = hm.sky_patch(**PATCH)
    cats = [mocks.survey_catalogue(ctx, patch, 20000, 200000, seed=300 + 7 * i, n_side=64, randoms_on_device=False)
            for i in range(6)]
    w0, w000 = measure_windows(ctx, cats[0], n, L)
    assert np.all(np.abs(w0) > 0) and np.all(np.abs(w000) > 0)
    c = cv.Covariance(ctx, cv.blocks_for(10, 5), cv.windows_for(w0, w000), keep_rows=8)
    meas = [measure_mock(ctx, cat, n, L) for cat in cats]
    for m in meas:
        c.push(m)
    X = np.array([c.concat(m) for m in meas]) / c.norm
    assert np.array_equal(c.vectors(), X)
    check_against_numpy(c, X)
    # the same through the reference's file layout
    root = str(tmp_path)
    rx = np.linspace(XI_KW["rmin"], XI_KW["rmax"], XI_KW["n_rbin"])
    rz = np.linspace(Z_KW["rmin"], Z_KW["rmax"], Z_KW["n_rbin"])
    from .test_covariance_files import write_xi, write_zeta
    for i, m in enumerate(meas):
        d = os.path.join(root, "mock", f"{i + 1:04d}")
        os.makedirs(d)
        for l in cv.XI_ELLS:
            write_xi(d, l, rx, m[cv.xi_name(l)])
        for mm in cv.ZETA_MULTIPOLES:
            write_zeta(d, *mm, rz, m[cv.zeta_name(*mm)])
    wd = os.path.join(root, "win")
    os.makedirs(wd)
    write_xi(wd, 0, rx, w0, suffix="_window")
    write_zeta(wd, 0, 0, 0, rz, w000, suffix="_window")
    cf, r2, r3 = cv.from_measurements(ctx, os.path.join(root, "mock"), wd, wd, 6, 10, 5, keep_rows=8)
    rd = lambda a: np.array(["%.7e" % v for v in np.ravel(a)], dtype=float)
    Xf = np.array([np.concatenate([rd(m[nm]) for nm in c.names]) for m in meas])
    Wf = np.concatenate([rd(w0)] * 3 + [rd(w000)] * 4)
    assert np.array_equal(cf.vectors(), Xf / Wf)
    check_against_numpy(cf, Xf / Wf)
    assert np.allclose(r2, rx) and r3.shape == (5, 5)
    cf.write(os.path.join(root, "out"), r2, r3)
    got = np.loadtxt(os.path.join(root, "out", "cov_xi0_zeta110.dat"))
    assert got.shape == (10, 25) and np.allclose(got, cf.block("xi0", "zeta110"), rtol=1e-15, atol=0)
    mean_file = np.loadtxt(os.path.join(root, "out", "zeta202_mean.dat"))
    assert mean_file.shape == (25, 4)
    c.close()
    cf.close()
