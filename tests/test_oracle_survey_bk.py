"""Pins the oracle's survey-geometry bispectrum and three-point functions
(bispec.hpp:4-874) independently of its own code paths:

  - bk_save equals the direct Fourier-space triangle sum
    sum_{m1 m2 M} w sum_{q1 in S1, q2 in S2} Y1(q1) a(q1) Y2(q2) a(q2) a_LM(-q1-q2),
    a = FFT(w (D - alpha R)) / W, a_LM the Y_LM(los)-weighted field / W;
  - sn_save of B_000 equals N + P(k1) + P(k2) + the ijk sweep built with
    numpy.fft + scipy j0 from the w^2-weighted D + alpha^2 R field and
    sum w_D^3 - alpha^3 sum w_R^3 (field.hpp:249-287, 1220-1240);
  - zeta_save (survey and window) equals the Fourier double sum, and the
    window zeta shot noise the numpy transform with sum_R w^2.
"""
import numpy as np
import pytest
from scipy.special import spherical_jn

from oracle import pyoracle as po

from .test_oracle_zeta import weights, ylm


def survey_catalogues(seed, L):
    rng = np.random.default_rng(seed)
    lo = np.array([300.0, 450.0, 600.0])
    size = 0.8 * L
    c = lo + rng.random((6, 3)) * size
    nD, nR = 1200, 4000
    a = np.clip(c[rng.integers(0, 6, nD // 2)] + rng.normal(0.0, 0.05 * size, (nD // 2, 3)), lo, lo + size)
    posD = np.concatenate([a, lo + rng.random((nD - nD // 2, 3)) * size])
    posR = lo + rng.random((nR, 3)) * size
    wD = 1.0 / (1.0 + 0.4 * rng.random(nD))
    wR = 1.0 / (1.0 + 0.4 * rng.random(nR))
    return posD, wD, posR, wR


def grid_k(n, L):
    idx = np.arange(n)
    sgn = np.where(idx < n // 2, idx, idx - n)
    G = np.stack(np.meshgrid(sgn, sgn, sgn, indexing="ij"), -1).reshape(-1, 3)
    return G, G * (2 * np.pi / L)


def weighted_field(p, pos, los, w, L, M, k, conj):
    y = np.array([ylm(L, M, v) for v in los])
    if conj:
        y = np.conj(y)
    wt = np.stack([(y * w ** k).real, (y * w ** k).imag], -1)
    return po.calc_field(p, pos, wt)


def all_weights(l1, l2, L):
    out = []
    for m1 in range(-l1, l1 + 1):
        for m2 in range(-l2, l2 + 1):
            for M in range(-L, L + 1):
                w = po.wigner3j(l1, l2, L, 0, 0, 0) * po.wigner3j(l1, l2, L, m1, m2, M)
                w *= (2 * L + 1) * (2 * l1 + 1) * (2 * l2 + 1)
                if abs(w) >= 1e-10:
                    out.append((m1, m2, M, w))
    return out


@pytest.mark.parametrize("ells", [(0, 0, 0), (2, 0, 2), (1, 1, 0)])
def test_survey_bk_bruteforce_triangles(ells):
    l1, l2, L_ = ells
    n, L = 10, 900.0
    p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.04, n_kbin=3, ith_kbin=1, ell1=l1, ell2=l2, ELL=L_)
    kb = po.kbin(p)
    s = po.Survey(p, *survey_catalogues(3, L))
    _, bk_save, _ = s.bk(p)
    G, K = grid_k(n, L)
    kmag = np.linalg.norm(K, axis=1)
    W = np.array([po.assignment_function(p, k) for k in K])
    a00 = po.fourier_transform(p, s.field(p, 0, 0)).ravel() / W
    dkb = kb[1] - kb[0]

    def shell(kc):
        lo = kc - dkb * 0.5 if kc > dkb * 0.5 else 0.0
        return np.nonzero((kmag > lo) & (kmag <= kc + dkb * 0.5))[0]

    lin = lambda t: (t[..., 0] * n + t[..., 1]) * n + t[..., 2]
    s1 = shell(kb[1])
    expect = np.zeros(len(kb), complex)
    for m1, m2, M, w in all_weights(l1, l2, L_):
        aLM = po.fourier_transform(p, s.field(p, L_, M)).ravel() / W
        Y1 = np.array([ylm(l1, m1, k) for k in K])
        Y2 = np.array([ylm(l2, m2, k) for k in K])
        for ik in range(len(kb)):
            s2 = shell(kb[ik])
            q3 = lin((-(G[s1][:, None, :] + G[s2][None, :, :])) % n)
            tot = np.sum((Y1 * a00)[s1][:, None] * (Y2 * a00)[s2][None, :] * aLM[q3])
            expect[ik] += w * tot / (len(s1) * len(s2) * L ** 3) * (L ** 3 / n ** 3) * n ** 3
    assert np.abs(bk_save - expect).max() <= 1e-10 * np.abs(expect).max()


def test_survey_bk000_shotnoise_numpy():
    n, L = 10, 900.0
    p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.04, n_kbin=3, ith_kbin=2)
    kb = po.kbin(p)
    s = po.Survey(p, *survey_catalogues(5, L))
    _, _, sn_save = s.bk(p)
    d00 = po.fourier_transform(p, s.field(p, 0, 0))
    N00 = po.fourier_transform(p, weighted_field(p, s.posD, s.losD, s.wD, 0, 0, 2, True)
                               + s.alpha ** 2 * weighted_field(p, s.posR, s.losR, s.wR, 0, 0, 2, True))
    sn = np.sum(s.wD ** 3) - s.alpha ** 3 * np.sum(s.wR ** 3)
    pk, _ = po.calc_power_spectrum(p, d00, N00, kb, sn, 0, 0)
    G, K = grid_k(n, L)
    S = np.array([po.shotnoise_function(p, k) for k in K]).reshape(n, n, n)
    W = np.array([po.assignment_function(p, k) for k in K]).reshape(n, n, n)
    g = (d00 * np.conj(N00) - sn * S) / W ** 2 / L ** 3
    xi = np.fft.ifftn(g) * g.size
    r = np.linalg.norm(G * (L / n), axis=1).reshape(n, n, n)
    for ik in range(len(kb)):
        ijk = np.sum(spherical_jn(0, kb[2] * r) * spherical_jn(0, kb[ik] * r) * xi) * (L ** 3 / n ** 3)
        expect = sn + pk[2] + pk[ik] + ijk
        assert abs(sn_save[ik] - expect) <= 1e-9 * abs(expect)


@pytest.mark.parametrize("window", [False, True])
@pytest.mark.parametrize("ells", [(0, 0, 0), (1, 1, 0)])
def test_survey_zeta_triple_part_fourier_double_sum(ells, window):
    l1, l2, L_ = ells
    n, L = 8, 900.0
    p = po.param(n, L, po.TSC, rmin=150.0, rmax=350.0, n_rbin=3, ith_rbin=1, ell1=l1, ell2=l2, ELL=L_)
    rb = po.rbin(p)
    s = po.Survey(p, *survey_catalogues(7, L))
    _, zeta_save, _ = s.zeta(p, window=window)
    G, K = grid_k(n, L)
    kmag = np.linalg.norm(K, axis=1)
    W = np.array([po.assignment_function(p, k) for k in K])
    if window:
        f00 = weighted_field(p, s.posR, s.losR, s.wR, 0, 0, 1, False)
    else:
        f00 = s.field(p, 0, 0)
    a = po.fourier_transform(p, f00).ravel() / W
    lin = lambda t: (t[..., 0] * n + t[..., 1]) * n + t[..., 2]
    q3 = lin((-(G[:, None, :] + G[None, :, :])) % n)
    expect = np.zeros(len(rb), complex)
    for m1, m2, M, w in all_weights(l1, l2, L_):
        fLM = weighted_field(p, s.posR, s.losR, s.wR, L_, M, 1, False) if window else s.field(p, L_, M)
        a3 = (po.fourier_transform(p, fLM).ravel() / W)[q3]
        Y1 = np.array([ylm(l1, m1, k) for k in K])
        Y2 = np.array([ylm(l2, m2, k) for k in K])
        A1 = spherical_jn(l1, kmag * rb[1]) * Y1 * a
        for ir, r2 in enumerate(rb):
            A2 = spherical_jn(l2, kmag * r2) * Y2 * a
            expect[ir] += w * (1j ** (l1 + l2)) * np.einsum("i,j,ij->", A1, A2, a3) / L ** 6
    assert np.abs(zeta_save - expect).max() <= 1e-8 * np.abs(expect).max()


def test_survey_window_zeta_shotnoise_numpy():
    n, L = 10, 900.0
    rmin, rmax, nr, ith = 120.0, 360.0, 4, 1
    p = po.param(n, L, po.TSC, rmin=rmin, rmax=rmax, n_rbin=nr, ith_rbin=ith)
    rb = po.rbin(p)
    s = po.Survey(p, *survey_catalogues(11, L))
    _, _, sn_save = s.zeta(p, window=True)
    d = po.fourier_transform(p, weighted_field(p, s.posR, s.losR, s.wR, 0, 0, 1, False))
    N = po.fourier_transform(p, weighted_field(p, s.posR, s.losR, s.wR, 0, 0, 2, True))
    sn = np.sum(s.wR ** 2)
    G, K = grid_k(n, L)
    S = np.array([po.shotnoise_function(p, k) for k in K]).reshape(n, n, n)
    W = np.array([po.assignment_function(p, k) for k in K]).reshape(n, n, n)
    xi = (np.fft.ifftn((d * np.conj(N) - sn * S) / W ** 2 / L ** 3) * n ** 3).ravel()
    r = np.linalg.norm(G * (L / n), axis=1)
    centre = (r / 0.5 + 0.5).astype(int) * 0.5
    drb = np.diff(rb)
    lower, upper = rb[ith] - drb[ith - 1] * 0.5, rb[ith] + drb[ith] * 0.5
    sel = (centre > lower) & (centre <= upper)
    expect = np.sum(xi[sel]) / sel.sum() ** 2 / (L ** 3 / n ** 3)
    assert abs(sn_save[ith] - expect) <= 1e-10 * abs(expect)
