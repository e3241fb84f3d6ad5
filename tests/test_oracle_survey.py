"""Pins the oracle's survey restatement (main.cpp:109-166, particle.hpp,
field.hpp:186-228, 449-503, 1164-1188, powerspec.hpp:4-91):

  - Vsurvey, alpha, nbar and the LOS against closed forms (randoms on the
    nodes of a sub-block with NGP give Vsurvey = m^3 dV exactly);
  - the survey P_l against a direct Yamamoto sum: by the addition theorem the
    reference's sum over M, m1 with 3j weights collapses to
    (2l+1) L_l(k.los) per particle, evaluated here particle by particle with
    the stencil phases sum_c W_pc exp(-i k.x_c), independently of the
    Y_lm-weighted meshes and the 3j weights.
"""
import numpy as np
import pytest
from scipy.special import eval_legendre

from oracle import pyoracle as po


def survey_catalogue(nD, nR, lo, size, seed):
    rng = np.random.default_rng(seed)
    posD = lo + rng.random((nD, 3)) * size
    posR = lo + rng.random((nR, 3)) * size
    wD = 0.5 + rng.random(nD)
    wR = 0.5 + rng.random(nR)
    return posD, wD, posR, wR


def test_survey_volume_alpha_nbar_los_closed_forms():
    n, L, m = 16, 800.0, 10
    dx = L / n
    p = po.param(n, L, po.NGP)
    i0 = np.array([7, 11, 13])
    g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)
    posR = (g + i0) * dx
    rng = np.random.default_rng(1)
    posD = (i0 + rng.random((200, 3)) * (m - 1)) * dx
    wD = 0.5 + rng.random(200)
    s = po.Survey(p, posD, wD, posR, np.ones(len(posR)))
    assert s.vsurvey == pytest.approx(m ** 3 * dx ** 3, rel=1e-12)
    assert s.alpha == pytest.approx(wD.sum() / len(posR), rel=1e-14)
    assert s.nbar == pytest.approx(wD.sum() ** 2 / s.vsurvey / (wD ** 2).sum(), rel=1e-14)
    assert np.allclose(s.shift, i0 * dx - dx, rtol=1e-14)
    # LOS of the unshifted positions
    assert np.allclose(s.losD, posD / np.linalg.norm(posD, axis=1)[:, None], rtol=1e-14)
    assert np.allclose(s.posD, posD - s.shift, rtol=1e-14)


def tsc_stencil(p, pos):
    """Per-particle (27 linear cells, 27 weights) by an independent vectorised TSC."""
    n = np.array(p.n_mesh[:])
    L = np.array(p.boxsize[:])
    xp = pos * n / L
    i1 = np.floor(xp + 0.5).astype(int)
    d = xp - i1
    ws = np.stack([0.5 * (0.5 - d) ** 2, 0.75 - d ** 2, 0.5 * (0.5 + d) ** 2])  # (3, np, 3)
    cells, wts = [], []
    for a in range(3):
        for b in range(3):
            for c in range(3):
                ix, iy, iz = i1[:, 0] + a - 1, i1[:, 1] + b - 1, i1[:, 2] + c - 1
                cells.append((ix * n[1] + iy) * n[2] + iz)
                wts.append(ws[a, :, 0] * ws[b, :, 1] * ws[c, :, 2])
    return np.stack(cells, 1), np.stack(wts, 1)


@pytest.mark.parametrize("ell", [0, 1, 2])
def test_survey_pk_equals_direct_yamamoto_sum(ell):
    n, L = 12, 600.0
    p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.1, n_kbin=5, ell1=ell, ELL=ell)
    posD, wD, posR, wR = survey_catalogue(150, 600, np.array([400.0, 700.0, 900.0]), 400.0, seed=ell + 3)
    s = po.Survey(p, posD, wD, posR, wR)
    got, nm = s.pk(p)

    idx = np.arange(n)
    sgn = np.where(idx < n // 2, idx, idx - n)
    dk = 2 * np.pi / L
    K = np.stack(np.meshgrid(sgn, sgn, sgn, indexing="ij"), -1).reshape(-1, 3) * dk
    kmag = np.linalg.norm(K, axis=1)
    khat = np.where(kmag[:, None] > 0, K / np.where(kmag > 0, kmag, 1)[:, None], 0.0)
    dx = L / n

    def stencil_phase(pos):
        cells, wts = tsc_stencil(p, pos)
        assert np.all((cells >= 0) & (cells < n ** 3))  # the catalogue sits inside the mesh
        ci = np.stack(np.unravel_index(cells, (n, n, n)), -1) * dx  # (np, 27, 3)
        ph = np.exp(-1j * np.einsum("kc,pqc->pqk", K, ci))            # (np, 27, nk)
        return np.einsum("pq,pqk->pk", wts, ph)

    def legendre(los):
        return eval_legendre(ell, khat @ los.T).T if ell else np.ones((len(los), len(K)))

    SD, SR = stencil_phase(s.posD), stencil_phase(s.posR)
    LD, LR = legendre(s.losD), legendre(s.losR)
    A = (wD[:, None] * LD * SD).sum(0) - s.alpha * (wR[:, None] * LR * SR).sum(0)
    F00 = (wD[:, None] * SD).sum(0) - s.alpha * (wR[:, None] * SR).sum(0)
    SN = (wD[:, None] ** 2 * LD).sum(0) + s.alpha ** 2 * (wR[:, None] ** 2 * LR).sum(0)
    if ell:
        SN = np.where(kmag > 0, SN, 0.0)
        A = np.where(kmag > 0, A, 0.0)
    W = np.array([po.assignment_function(p, k) for k in K])
    S = np.array([po.shotnoise_function(p, k) for k in K])
    val = (2 * ell + 1) * (A * np.conj(F00) - SN * S) / W ** 2
    centre = (kmag / 1e-4 + 0.5).astype(int) * 1e-4
    kb = po.kbin(p)
    dkb = kb[1] - kb[0]
    norm = s.vsurvey / wD.sum() ** 2
    for j, c in enumerate(kb):
        lo = c - dkb * 0.5 if c > dkb * 0.5 else 0.0
        sel = (centre > lo) & (centre <= c + dkb * 0.5)
        assert nm[j] == sel.sum()
        expect = norm * val[sel].sum() / sel.sum()
        assert abs(got[j] - expect) <= 1e-10 * abs(norm * val[sel]).mean() + 1e-300


@pytest.mark.parametrize("ell,window", [(0, False), (2, False), (1, False), (2, True)])
def test_survey_xi_equals_direct_sum(ell, window):
    """powerspec.hpp:93-270: sum_M w_M Y_{l,-M}(r) IFFT[(F_LM conj F_00 - SN_M S)/W^2/V]
    equals, by the addition theorem, sum_k e^{ik.r} [(2l+1) A_l(r, k) conj F_00(k)
    - (2l+1) SN_l(r) S(k)] / W(k)^2 / V with A_l(r, k) = sum_p w_p L_l(r.los_p) S_p(k),
    evaluated here point by point (window: the random catalogue alone)."""
    n, L = 10, 500.0
    p = po.param(n, L, po.TSC, rmin=40.0, rmax=240.0, n_rbin=6, ell1=ell, ELL=ell)
    posD, wD, posR, wR = survey_catalogue(100, 400, np.array([300.0, -700.0, 500.0]), 330.0, seed=11 + ell)
    s = po.Survey(p, posD, wD, posR, wR)
    got, nm = s.xi(p, window=window)

    idx = np.arange(n)
    sgn = np.where(idx < n // 2, idx, idx - n)
    grid = np.stack(np.meshgrid(sgn, sgn, sgn, indexing="ij"), -1).reshape(-1, 3)
    K = grid * (2 * np.pi / L)
    Rv = grid * (L / n)
    rmag = np.linalg.norm(Rv, axis=1)
    rhat = np.where(rmag[:, None] > 0, Rv / np.where(rmag > 0, rmag, 1)[:, None], 0.0)
    dx = L / n

    def stencil_phase(pos):
        cells, wts = tsc_stencil(p, pos)
        assert np.all((cells >= 0) & (cells < n ** 3))
        ci = np.stack(np.unravel_index(cells, (n, n, n)), -1) * dx
        ph = np.exp(-1j * np.einsum("kc,pqc->pqk", K, ci))
        return np.einsum("pq,pqk->pk", wts, ph)

    def leg(los):
        out = eval_legendre(ell, rhat @ los.T) if ell else np.ones((len(Rv), len(los)))
        if ell:
            out[rmag == 0] = 0.0
        return out  # (N_r, N_p)

    SR, LR = stencil_phase(s.posR), leg(s.losR)
    if window:
        A = (LR * wR[None, :]) @ SR
        F00 = wR @ SR
        SN = (LR * (wR ** 2)[None, :]).sum(1)
        sw = wR.sum()
    else:
        SD, LD = stencil_phase(s.posD), leg(s.losD)
        A = (LD * wD[None, :]) @ SD - s.alpha * ((LR * wR[None, :]) @ SR)
        F00 = wD @ SD - s.alpha * (wR @ SR)
        SN = (LD * (wD ** 2)[None, :]).sum(1) + s.alpha ** 2 * (LR * (wR ** 2)[None, :]).sum(1)
        sw = wD.sum()
    W = np.array([po.assignment_function(p, k) for k in K])
    S = np.array([po.shotnoise_function(p, k) for k in K])
    g = (2 * ell + 1) * (A * np.conj(F00)[None, :] - SN[:, None] * S[None, :]) / W[None, :] ** 2 / L ** 3
    xi = (g * np.exp(1j * (Rv @ K.T))).sum(1)
    centre = (rmag / 0.5 + 0.5).astype(int) * 0.5
    rb = po.rbin(p)
    drb = rb[1] - rb[0]
    norm = s.vsurvey / sw ** 2
    for j, c in enumerate(rb):
        lo = c - drb * 0.5 if c > drb * 0.5 else 0.0
        sel = (centre > lo) & (centre <= c + drb * 0.5)
        assert nm[j] == sel.sum()
        if sel.sum() == 0:
            continue
        expect = norm * xi[sel].sum() / sel.sum()
        assert abs(got[j] - expect) <= 1e-10 * norm * np.abs(xi[sel]).mean()
