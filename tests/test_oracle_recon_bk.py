"""Pins the oracle's reconstructed-box bispectrum and three-point function
(bispec.hpp:1426-1753, 1755-1975) independently of its own code paths:

  - bk_save equals the direct Fourier-space triangle sum over the shells of
    d = FFT(D - alpha R);
  - sn_save equals numpy.fft + scipy j0 with the cross spectrum of
    D - alpha R and D + alpha^2 R (field.hpp:528-605) and the shot noise
    N_D - alpha^3 N_R (field.hpp:1242-1250), real and imaginary parts;
  - zeta_save equals the Fourier double sum with d, and the zeta shot noise
    the numpy transform of the same cross spectrum, binned with the 3PCF edges.
"""
import numpy as np
import pytest
from scipy.special import spherical_jn

from oracle import pyoracle as po

from .conftest import clustered_catalogue, uniform_catalogue
from .test_oracle_zeta import weights, ylm


def recon_fields(p, D, R, alpha):
    fD = po.calc_field(p, D)
    fR = po.calc_field(p, R)
    dn = po.fourier_transform(p, fD - alpha * fR)
    nn = po.fourier_transform(p, fD + alpha ** 2 * fR)
    return dn, nn


def catalogues(L, seed):
    D = clustered_catalogue(1500, L, seed=seed, nclump=6, sigma=40.0)
    R = uniform_catalogue(4000, L, seed=seed + 1)
    return D, R, len(D) / len(R)


def grid_k(n, L):
    idx = np.arange(n)
    sgn = np.where(idx < n // 2, idx, idx - n)
    G = np.stack(np.meshgrid(sgn, sgn, sgn, indexing="ij"), -1).reshape(-1, 3)
    return sgn, G, G * (2 * np.pi / L)


def test_recon_bk_bruteforce_triangles():
    n, L = 12, 600.0
    p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.05, n_kbin=4, ith_kbin=1)
    kb = po.kbin(p)
    D, R, alpha = catalogues(L, 41)
    _, bk_save, _ = po.box_recon_bk(p, D, R, alpha)
    dn, _ = recon_fields(p, D, R, alpha)
    sgn, G, K = grid_k(n, L)
    kmag = np.linalg.norm(K, axis=1)
    W = np.array([po.assignment_function(p, k) for k in K])
    a = dn.ravel() / W
    dkb = kb[1] - kb[0]

    def shell(kc):
        lo = kc - dkb * 0.5 if kc > dkb * 0.5 else 0.0
        return np.nonzero((kmag > lo) & (kmag <= kc + dkb * 0.5))[0]

    lin = lambda t: (t[..., 0] * n + t[..., 1]) * n + t[..., 2]
    s1 = shell(kb[1])
    for ik in range(len(kb)):
        s2 = shell(kb[ik])
        q3 = lin((-(G[s1][:, None, :] + G[s2][None, :, :])) % n)
        tot = np.sum(a[s1][:, None] * a[s2][None, :] * a[q3])
        expect = tot / (len(s1) * len(s2) * L ** 3) * (L ** 3 / n ** 3) * n ** 3
        assert abs(bk_save[ik] - expect) <= 1e-10 * abs(expect)


def test_recon_bk_shotnoise_numpy():
    n, L = 12, 600.0
    p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.05, n_kbin=4, ith_kbin=2)
    kb = po.kbin(p)
    D, R, alpha = catalogues(L, 43)
    _, _, sn_save = po.box_recon_bk(p, D, R, alpha)
    dn, nn = recon_fields(p, D, R, alpha)
    sn = len(D) - alpha ** 3 * len(R)
    pk, _ = po.calc_power_spectrum(p, dn, nn, kb, float(sn), 0, 0)
    sgn, G, K = grid_k(n, L)
    S = np.array([po.shotnoise_function(p, k) for k in K]).reshape(n, n, n)
    W = np.array([po.assignment_function(p, k) for k in K]).reshape(n, n, n)
    g = (dn * np.conj(nn) - sn * S) / W ** 2 / L ** 3
    xi = np.fft.ifftn(g) * g.size
    r = np.linalg.norm(G * (L / n), axis=1).reshape(n, n, n)
    assert np.abs(pk.imag).max() > 1e-12 * np.abs(pk.real).max()  # the cross spectrum is complex
    for ik in range(len(kb)):
        ijk = np.sum(spherical_jn(0, kb[2] * r) * spherical_jn(0, kb[ik] * r) * xi) * (L ** 3 / n ** 3)
        expect = sn + pk[2] + pk[ik] + ijk
        assert abs(sn_save[ik] - expect) <= 1e-9 * abs(expect)


@pytest.mark.parametrize("ells", [(0, 0, 0), (1, 1, 0)])
def test_recon_zeta_triple_part_fourier_double_sum(ells):
    l1, l2, L_ = ells
    n, L = 8, 400.0
    p = po.param(n, L, po.TSC, rmin=50.0, rmax=150.0, n_rbin=3, ith_rbin=1, ell1=l1, ell2=l2, ELL=L_)
    rb = po.rbin(p)
    D, R, alpha = catalogues(L, 47)
    _, zeta_save, _ = po.box_recon_zeta(p, D, R, alpha)
    dn, _ = recon_fields(p, D, R, alpha)
    sgn, G, K = grid_k(n, L)
    kmag = np.linalg.norm(K, axis=1)
    W = np.array([po.assignment_function(p, k) for k in K])
    a = dn.ravel() / W
    lin = lambda t: (t[..., 0] * n + t[..., 1]) * n + t[..., 2]
    a3 = a[lin((-(G[:, None, :] + G[None, :, :])) % n)]
    expect = np.zeros(len(rb), complex)
    for m1, m2, w in weights(l1, l2, L_):
        Y1 = np.array([ylm(l1, m1, k) for k in K])
        Y2 = np.array([ylm(l2, m2, k) for k in K])
        A1 = spherical_jn(l1, kmag * rb[1]) * Y1 * a
        for ir, r2 in enumerate(rb):
            A2 = spherical_jn(l2, kmag * r2) * Y2 * a
            expect[ir] += w * (1j ** (l1 + l2)) * np.einsum("i,j,ij->", A1, A2, a3) / L ** 6
    assert np.abs(zeta_save - expect).max() <= 1e-8 * np.abs(expect).max()


@pytest.mark.parametrize("ells", [(0, 0, 0), (1, 1, 0)])
def test_recon_zeta_shotnoise_numpy(ells):
    l1, l2, L_ = ells
    n, L = 10, 500.0
    rmin, rmax, nr, ith = 40.0, 160.0, 4, 2
    p = po.param(n, L, po.TSC, rmin=rmin, rmax=rmax, n_rbin=nr, ith_rbin=ith, ell1=l1, ell2=l2, ELL=L_)
    rb = po.rbin(p)
    D, R, alpha = catalogues(L, 53)
    _, _, sn_save = po.box_recon_zeta(p, D, R, alpha)
    dn, nn = recon_fields(p, D, R, alpha)
    sn = len(D) - alpha ** 3 * len(R)
    sgn, G, K = grid_k(n, L)
    S = np.array([po.shotnoise_function(p, k) for k in K]).reshape(n, n, n)
    W = np.array([po.assignment_function(p, k) for k in K]).reshape(n, n, n)
    g = (dn * np.conj(nn) - sn * S) / W ** 2 / L ** 3
    xi = (np.fft.ifftn(g) * g.size).ravel()
    Rv = G * (L / n)
    r = np.linalg.norm(Rv, axis=1)
    centre = (r / 0.5 + 0.5).astype(int) * 0.5
    drb = np.diff(rb)
    lower = rb[ith] - drb[ith - 1] * 0.5
    upper = rb[ith] + drb[ith] * 0.5
    sel = (centre > lower) & (centre <= upper)
    expect = 0.0
    for m1, m2, w in weights(l1, l2, L_):
        Y = np.array([ylm(l1, m1, v) * ylm(l2, m2, v) for v in Rv])
        expect += w * np.sum(xi[sel] * Y[sel]) * (-1.0) ** (l1 + l2) / sel.sum() ** 2 / (L ** 3 / n ** 3)
    assert abs(sn_save[ith] - expect) <= 1e-10 * abs(expect)
    assert abs(expect.imag) > 1e-8 * abs(expect.real)  # the cross spectrum gives an imaginary part
