"""Pins the oracle's box 3PCF (bispec.hpp:1205-1424, field.hpp:403-445 and
979-1113):

  - zeta_save against the Fourier-space double sum it equals,
    i^(l1+l2) / V^2 sum_m w_m sum_{k1,k2} j_l1(k1 r1) Y1(k1) a(k1)
    j_l2(k2 r2) Y2(k2) a(k2) a(-k1-k2), a = delta/W, with scipy's j_l;
  - sn_save against numpy.fft + the reference's 3PCF bin edges and the
    (-1)^(l1+l2) / n_mode^2 / dV scaling (Q14), on the ith_rbin bin only.
"""
import numpy as np
import pytest
from scipy.special import spherical_jn

from oracle import pyoracle as po

from .conftest import clustered_catalogue


def ylm(l, m, v):
    out = np.zeros(2)
    po.lib().or_ylm(l, m, po.dp(np.ascontiguousarray(v, dtype=np.float64)), po.dp(out))
    return out[0] + 1j * out[1]


def weights(l1, l2, L):
    out = []
    for m1 in range(-l1, l1 + 1):
        m2 = -m1
        if abs(m2) > l2:
            continue
        w = po.wigner3j(l1, l2, L, 0, 0, 0) * po.wigner3j(l1, l2, L, m1, m2, 0)
        w *= (2 * L + 1) * (2 * l1 + 1) * (2 * l2 + 1)
        if abs(w) >= 1e-10:
            out.append((m1, m2, w))
    return out


@pytest.mark.parametrize("ells", [(0, 0, 0), (1, 1, 0), (2, 0, 2), (1, 1, 2)])
def test_zeta_triple_part_fourier_double_sum(ells):
    l1, l2, L_ = ells
    n, L, N = 8, 400.0, 3000
    p = po.param(n, L, po.TSC, rmin=50.0, rmax=150.0, n_rbin=3, ith_rbin=1, ell1=l1, ell2=l2, ELL=L_)
    rb = po.rbin(p)
    pos = clustered_catalogue(N, L, seed=23, nclump=5, sigma=30.0)
    _, zeta_save, _ = po.box_zeta(p, pos)
    d = po.fourier_transform(p, po.box_density_fluctuation(p, pos))
    idx = np.arange(n)
    sgn = np.where(idx < n // 2, idx, idx - n)
    g = np.stack(np.meshgrid(idx, idx, idx, indexing="ij"), -1).reshape(-1, 3)
    K = sgn[g] * (2 * np.pi / L)
    kmag = np.linalg.norm(K, axis=1)
    W = np.array([po.assignment_function(p, k) for k in K])
    a = d.ravel() / W
    neg = ((-g) % n)
    lin = lambda t: (t[..., 0] * n + t[..., 1]) * n + t[..., 2]
    a3 = a[lin((-(g[:, None, :] + g[None, :, :])) % n)]  # a(-k1-k2)
    expect = np.zeros(len(rb), complex)
    for m1, m2, w in weights(l1, l2, L_):
        Y1 = np.array([ylm(l1, m1, k) for k in K])
        Y2 = np.array([ylm(l2, m2, k) for k in K])
        A1 = spherical_jn(l1, kmag * rb[1]) * Y1 * a
        for ir, r2 in enumerate(rb):
            A2 = spherical_jn(l2, kmag * r2) * Y2 * a
            expect[ir] += w * (1j ** (l1 + l2)) * np.einsum("i,j,ij->", A1, A2, a3) / L ** 6
    assert np.abs(zeta_save - expect).max() <= 1e-8 * np.abs(expect).max()
    assert neg.shape == g.shape


@pytest.mark.parametrize("ells", [(0, 0, 0), (1, 1, 0), (1, 0, 1)])
def test_zeta_shotnoise_numpy(ells):
    l1, l2, L_ = ells
    n, L, N = 10, 500.0, 2000
    rmin, rmax, nr, ith = 40.0, 160.0, 4, 2
    p = po.param(n, L, po.TSC, rmin=rmin, rmax=rmax, n_rbin=nr, ith_rbin=ith, ell1=l1, ell2=l2, ELL=L_)
    rb = po.rbin(p)
    pos = clustered_catalogue(N, L, seed=29, nclump=5, sigma=40.0)
    _, _, sn_save = po.box_zeta(p, pos)
    d = po.fourier_transform(p, po.box_density_fluctuation(p, pos))
    nn = d.copy()
    nn[0, 0, 0] += N
    idx = np.arange(n)
    sgn = np.where(idx < n // 2, idx, idx - n)
    G = np.stack(np.meshgrid(sgn, sgn, sgn, indexing="ij"), -1).reshape(-1, 3)
    K = G * (2 * np.pi / L)
    S = np.array([po.shotnoise_function(p, k) for k in K]).reshape(n, n, n)
    W = np.array([po.assignment_function(p, k) for k in K]).reshape(n, n, n)
    g = (d * np.conj(nn) - N * S) / W ** 2 / L ** 3
    xi = (np.fft.ifftn(g) * g.size).ravel()
    R = G * (L / n)
    r = np.linalg.norm(R, axis=1)
    centre = (r / 0.5 + 0.5).astype(int) * 0.5
    drb = np.diff(rb)
    lower = rb[ith] - drb[ith - 1] * 0.5 if ith > 0 else max(rb[0] - drb[0] * 0.5, 0.0)
    upper = rb[ith] + drb[ith - 1] * 0.5 if ith == nr - 1 else rb[ith] + drb[ith] * 0.5
    sel = (centre > lower) & (centre <= upper)
    expect = 0.0
    for m1, m2, w in weights(l1, l2, L_):
        Y = np.array([ylm(l1, m1, v) * ylm(l2, m2, v) for v in R])
        expect += w * np.sum(xi[sel] * Y[sel]) * (-1.0) ** (l1 + l2) / sel.sum() ** 2 / (L ** 3 / n ** 3)
    assert abs(sn_save[ith] - expect) <= 1e-10 * abs(expect)
    others = np.delete(np.arange(nr), ith)
    assert np.all(sn_save[others] == 0)
