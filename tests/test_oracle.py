"""Pins the FP64 CPU oracle (oracle/hm_oracle.c) before anything is compared
against it.

What pins what (reference file:line of the restated code in brackets):
  - FFT [fftw_plan_dft_3d, field.hpp:321/339]: numpy.fft fixtures (golden),
    including prime and non-power-of-two lengths; Parseval; impulse.
  - Wigner 3j [common.hpp:29]: sympy exact values (golden).
  - sphPlm and reduced Y_lm [tools.hpp:12-68]: sympy assoc_legendre / Ynm (golden).
  - j_l and its spline [bessel.hpp:10-37]: mpmath (golden); spline vs closed form.
  - mass assignment [field.hpp:145-184]: single particle on a node gives
    (1/8, 3/4, 1/8)^3; mass conservation; hand-derived Q1 spill/drop coords;
    an independent vectorised TSC.
  - window / shot-noise functions [field.hpp:700-717, 1293-1309]: hand values
    incl. the truncated-index quirk W(-n) = W(n-1).
  - P binning [field.hpp:765-852]: exact mode counts from a vectorised count;
    two-mode plane wave gives the closed-form bin value.
  - B [bispec.hpp:1100-1182]: brute-force triangle sum in Fourier space equals
    the FFT estimator; shot-noise ijk term equals a numpy/scipy evaluation.
"""
import json
import os

import numpy as np
import pytest
from scipy.special import spherical_jn

from oracle import pyoracle as po

from .conftest import clustered_catalogue

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def special():
    with open(os.path.join(GOLD, "special.json")) as f:
        return json.load(f)


def test_fft_golden():
    fx = np.load(os.path.join(GOLD, "fft_fixtures.npz"))
    for key in fx.files:
        if key.startswith("in_"):
            shape = tuple(int(s) for s in key[3:].split("x"))
            p = po.param(shape, (100.0, 110.0, 120.0))
            a = fx[key]
            f = po.fft3d(p, a, -1)
            b = po.fft3d(p, a, +1)
            ref_f, ref_b = fx["fwd_" + key[3:]], fx["bwd_" + key[3:]]
            assert np.abs(f - ref_f).max() <= 1e-12 * np.abs(ref_f).max()
            assert np.abs(b - ref_b).max() <= 1e-12 * np.abs(ref_b).max()
        if key.startswith("in1_"):
            n = key[4:]
            f = po.fft1d(fx[key], -1)
            assert np.abs(f - fx["fwd1_" + n]).max() <= 1e-12 * np.abs(fx["fwd1_" + n]).max()


def test_fft_impulse_and_parseval():
    p = po.param((8, 12, 16), 100.0)
    a = np.zeros((8, 12, 16), complex)
    a[3, 5, 7] = 2.0 - 1.0j
    f = po.fft3d(p, a, -1)
    i, j, k = np.meshgrid(np.arange(8), np.arange(12), np.arange(16), indexing="ij")
    expect = (2.0 - 1.0j) * np.exp(-2j * np.pi * (3 * i / 8 + 5 * j / 12 + 7 * k / 16))
    assert np.allclose(f, expect, atol=1e-13)
    rng = np.random.default_rng(1)
    b = rng.standard_normal((8, 12, 16)) + 1j * rng.standard_normal((8, 12, 16))
    fb = po.fft3d(p, b, -1)
    assert np.isclose(np.sum(np.abs(fb) ** 2), b.size * np.sum(np.abs(b) ** 2), rtol=1e-13)


def test_wigner3j_golden(special):
    for j1, j2, j3, m1, m2, m3, v in special["wigner3j"]:
        assert po.wigner3j(j1, j2, j3, m1, m2, m3) == pytest.approx(v, abs=1e-14)
    assert po.wigner3j(1, 1, 1, 0, 0, 0) == 0.0  # odd sum with zero m
    assert po.wigner3j(2, 2, 2, 1, 1, 1) == 0.0  # m1 + m2 + m3 != 0


def test_sphplm_and_ylm_golden(special):
    for l, m, x, v in special["sphPlm"]:
        assert po.sph_plm(l, m, x) == pytest.approx(v, rel=1e-12, abs=1e-14)
    for l, m, v, re, im in special["ylm_reduced"]:
        y = po.ylm(l, m, v)
        assert y.real == pytest.approx(re, abs=1e-12) and y.imag == pytest.approx(im, abs=1e-12), (l, m, v)
    # tools.hpp:18-20 and 32-34
    assert po.ylm(0, 0, [0.0, 0.0, 0.0]) == 1.0
    assert po.ylm(2, 1, [0.0, 0.0, 0.0]) == 0.0


def test_bessel_golden(special):
    for l, x, v in special["bessel_jl"]:
        got = po.bessel_jl(l, x)
        assert got == pytest.approx(v, rel=1e-11, abs=1e-15), (l, x)
    # the spline interpolates j_l to ~1e-10 away from the natural end condition
    for ell in (0, 1, 2):
        sp = po.BesselSpline(ell)
        for x in [0.0, 0.07, 0.5, 1.234, 17.5, 333.3, 9999.99]:
            assert sp(x) == pytest.approx(po.bessel_jl(ell, x), abs=2e-10)
        assert np.isnan(sp(10000.5))  # outside the spline domain (bessel.hpp:13)


def test_tsc_single_particle_on_node():
    L, n = 100.0, 16
    p = po.param(n, L, po.TSC)
    dx = L / n
    pos = np.array([[5 * dx, 7 * dx, 9 * dx]])
    f = po.calc_field(p, pos).real
    dV = dx ** 3
    w = np.array([0.125, 0.75, 0.125])
    expect = np.einsum("i,j,k->ijk", w, w, w) / dV
    assert np.allclose(f[4:7, 6:9, 8:11], expect, rtol=1e-14, atol=0)
    assert np.isclose(f.sum() * dV, 1.0, rtol=1e-14)


@pytest.mark.parametrize("assign", [po.NGP, po.CIC, po.TSC])
def test_mass_conservation_interior(assign):
    L, n = 200.0, 16
    p = po.param(n, L, assign)
    rng = np.random.default_rng(3)
    pos = 2 * L / n + rng.random((500, 3)) * (L - 4 * L / n)
    f = po.calc_field(p, pos)
    assert np.isclose(f.real.sum() * (L / n) ** 3, 500.0, rtol=1e-13)


def test_q1_unwrapped_stencil_coords():
    """field.hpp:175-176: a z index of n spills into (ix, iy + 1, 0); an x index
    of n or -1 is dropped, a y index of n spills into the next x plane."""
    L, n = 100.0, 8
    dx = L / n
    p = po.param(n, L, po.TSC)
    pos = np.array([[3.2 * dx, 4.1 * dx, (n - 0.3) * dx],    # iz1 = n -> stencil n-1, n, n+1
                    [(n - 0.2) * dx, 2.0 * dx, 2.0 * dx],    # ix1 = n -> n-1, n, n+1
                    [0.2 * dx, 2.0 * dx, 2.0 * dx]])         # ix1 = 0 -> -1, 0, 1
    c = po.cell_coords(p, pos)
    ix, iy = 3, 4
    got = set(c[0].tolist())
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for z in (n - 1, n, n + 1):
                lin = ((ix + di) * n + (iy + dj)) * n + z
                assert lin in got
    # z = n and n + 1 spill into the next row: (3, 4, n) is the linear index of (3, 5, 0)
    assert ((3 * n + 5) * n + 0) in got and ((3 * n + 5) * n + 1) in got
    # x = n + 1 row: dropped (all coords >= n^3 become -1)
    assert (c[1].reshape(3, 3, 3)[2] == -1).all()
    assert (c[1].reshape(3, 3, 3)[1] == -1).all()   # x = n
    assert (c[1].reshape(3, 3, 3)[0] >= 0).all()    # x = n - 1
    assert (c[2].reshape(3, 3, 3)[0] == -1).all()   # x = -1


def test_tsc_against_independent_vectorised():
    L, n = 300.0, 16
    p = po.param(n, L, po.TSC)
    rng = np.random.default_rng(5)
    pos = 2 * L / n + rng.random((300, 3)) * (L - 4 * L / n)
    f = po.calc_field(p, pos).real
    xp = pos * n / L
    i1 = np.floor(xp + 0.5).astype(int)
    d = xp - i1
    ws = [0.5 * (0.5 - d) ** 2, 0.75 - d ** 2, 0.5 * (0.5 + d) ** 2]
    g = np.zeros((n, n, n))
    for a in range(3):
        for b in range(3):
            for c in range(3):
                np.add.at(g, (i1[:, 0] + a - 1, i1[:, 1] + b - 1, i1[:, 2] + c - 1), ws[a][:, 0] * ws[b][:, 1] * ws[c][:, 2])
    assert np.allclose(f, g / (L / n) ** 3, rtol=1e-12, atol=1e-15)


def test_window_quirk_values():
    L, n = 1000.0, 32
    p = po.param(n, L, po.TSC)
    dk = 2 * np.pi / L
    sinc = lambda m: np.sin(np.pi * m / n) / (np.pi * m / n)
    assert po.assignment_function(p, [dk * 1, 0, 0]) == pytest.approx(sinc(1) ** 3, rel=1e-15)
    assert po.assignment_function(p, [-dk * 1, 0, 0]) == 1.0          # |n| - 1 = 0
    for m in (2, 5, 15):
        assert po.assignment_function(p, [0, -dk * m, 0]) == pytest.approx(sinc(m - 1) ** 3, rel=1e-15)
    assert po.assignment_function(p, [0, 0, -dk * 16]) == pytest.approx(sinc(15) ** 3, rel=1e-15)  # Nyquist
    s = np.sin(np.pi * 3 / n)
    assert po.shotnoise_function(p, [dk * 3, 0, 0]) == pytest.approx(1 - s * s + 2 * s ** 4 / 15, rel=1e-15)
    pc = po.param(n, L, po.CIC)
    assert po.assignment_function(pc, [dk * 2, dk, 0]) == pytest.approx((sinc(2) * sinc(1)) ** 2, rel=1e-15)
    assert po.shotnoise_function(pc, [0, -dk * 4, 0]) == pytest.approx(1 - 2 / 3 * np.sin(np.pi * 3 / n) ** 2,
                                                                        rel=1e-15)


def _mode_counts_numpy(n, L, kbin):
    """Fine histogram (1e-4) + coarse fold of field.hpp:793-837, vectorised."""
    dk = 2 * np.pi / L
    idx = np.arange(n)
    kv = np.where(idx < n // 2, idx, idx - n) * dk
    kx, ky, kz = np.meshgrid(kv, kv, kv, indexing="ij")
    kmag = np.sqrt(kx * kx + ky * ky + kz * kz)
    kindex = (kmag / 0.0001 + 0.5).astype(np.int64)
    hist = np.bincount(kindex[kindex < 100000].ravel(), minlength=100000)
    dkb = kbin[1] - kbin[0]
    centres = np.arange(100000) * 0.0001
    out = []
    for kb in kbin:
        lo = kb - dkb * 0.5 if kb > dkb * 0.5 else 0.0
        hi = kb + dkb * 0.5
        out.append(int(hist[(centres > lo) & (centres <= hi)].sum()))
    return np.array(out)


@pytest.mark.parametrize("n,L,kmin,kmax,nk", [(16, 500.0, 0.01, 0.2, 20), (24, 1000.0, 0.005, 0.195, 20)])
def test_pk_mode_counts(n, L, kmin, kmax, nk):
    p = po.param(n, L, po.TSC, kmin=kmin, kmax=kmax, n_kbin=nk)
    kb = po.kbin(p)
    f = np.ones((n, n, n), complex)
    _, nm = po.calc_power_spectrum(p, f, f, kb, 0.0, 0, 0)
    assert (nm == _mode_counts_numpy(n, L, kb)).all()


def _shell_and_pk_counts_numpy(n, L, kbin):
    """P bin counts (fine histogram + fold, field.hpp:793-837) and bispectrum shell counts
    (raw |k| in (max(k - dk/2, 0), k + dk/2], field.hpp:361-387), vectorised over x slabs."""
    n = list(n) if np.ndim(n) else [int(n)] * 3
    L = list(L) if np.ndim(L) else [float(L)] * 3
    kv = []
    for a in range(3):
        idx = np.arange(n[a])
        kv.append(np.where(idx < n[a] // 2, idx, idx - n[a]) * (2 * np.pi / L[a]))
    dkb = kbin[1] - kbin[0]
    lo = np.array([kb - dkb * 0.5 if kb > dkb * 0.5 else 0.0 for kb in kbin])
    hi = kbin + dkb * 0.5
    hist = np.zeros(100000, dtype=np.int64)
    shells = np.zeros(len(kbin), dtype=np.int64)
    ky, kz = np.meshgrid(kv[1], kv[2], indexing="ij")
    for kx in kv[0]:
        kmag = np.sqrt(kx * kx + ky * ky + kz * kz).ravel()
        kindex = (kmag / 0.0001 + 0.5).astype(np.int64)
        hist += np.bincount(kindex[kindex < 100000], minlength=100000)
        ks = np.sort(kmag)
        shells += np.searchsorted(ks, hi, side="right") - np.searchsorted(ks, lo, side="right")
    centres = np.arange(100000) * 0.0001
    pk = np.array([int(hist[(centres > lo[j]) & (centres <= hi[j])].sum()) for j in range(len(kbin))])
    return pk, shells


@pytest.mark.parametrize("n,L,kmin,kmax,nk", [(32, 1000.0, 0.02, 0.08, 4), (24, 700.0, 0.01, 0.2, 9),
                                               ((20, 36, 24), (1350.0, 2450.0, 1400.0), 0.01, 0.1, 10)])
def test_mode_counts_only(n, L, kmin, kmax, nk):
    """or_mode_counts equals the counts of calcPowerSpectrum and of every bispectrum shell of
    calcInverseFourierTransformForBispectrum, and an independent vectorised count."""
    p = po.param(n, L, po.TSC, kmin=kmin, kmax=kmax, n_kbin=nk)
    kb = po.kbin(p)
    pk, sh = po.mode_counts(p)
    f = np.ones(po.shape(p), complex)
    _, nm = po.calc_power_spectrum(p, f, f, kb, 0.0, 0, 0)
    assert np.array_equal(pk, nm)
    assert np.array_equal(sh, [po.ifft_for_bispectrum(p, f, k, kb[1] - kb[0])[1] for k in kb])
    ref_pk, ref_sh = _shell_and_pk_counts_numpy(n, L, kb)
    assert np.array_equal(pk, ref_pk) and np.array_equal(sh, ref_sh)


def test_mode_counts_metric_mesh():
    """512^3, L = 2000, 20 bins 0.015..0.205 (the bispectrum workload): the count-only oracle against the
    independent vectorised count."""
    p = po.param(512, 2000.0, po.TSC, kmin=0.015, kmax=0.205, n_kbin=20)
    pk, sh = po.mode_counts(p)
    ref_pk, ref_sh = _shell_and_pk_counts_numpy(512, 2000.0, po.kbin(p))
    assert np.array_equal(pk, ref_pk) and np.array_equal(sh, ref_sh)


def test_pk_plane_wave_closed_form():
    n, L = 16, 400.0
    p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.2, n_kbin=10)
    kb = po.kbin(p)
    f = np.zeros((n, n, n), complex)
    A = 3.0 + 4.0j
    f[2, 1, 0] = A
    f[-2, -1, 0] = np.conj(A)
    pk, nm = po.calc_power_spectrum(p, f, f, kb, 0.0, 0, 0)
    dk = 2 * np.pi / L
    kmag = np.sqrt(5.0) * dk
    W1 = po.assignment_function(p, [2 * dk, dk, 0.0])
    W2 = po.assignment_function(p, [-2 * dk, -dk, 0.0])
    j = int(np.argmin(np.abs(kb - kmag)))
    expect = (abs(A) ** 2 / W1 ** 2 + abs(A) ** 2 / W2 ** 2) / nm[j]
    assert pk[j].real == pytest.approx(expect, rel=1e-13)
    others = np.delete(np.arange(len(kb)), j)
    assert np.all(pk[others] == 0)


def test_box_bk_bruteforce_triangles():
    """bk_save of bispec.hpp:1159-1168 equals the direct Fourier-space triangle sum."""
    n, L = 12, 600.0
    p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.05, n_kbin=4, ith_kbin=1)
    kb = po.kbin(p)
    pos = clustered_catalogue(3000, L, seed=7, nclump=6, sigma=40.0)
    out, bk_save, sn_save = po.box_bk(p, pos)
    # brute force
    d = po.fourier_transform(p, po.box_density_fluctuation(p, pos))
    dk = 2 * np.pi / L
    idx = np.arange(n)
    sgn = np.where(idx < n // 2, idx, idx - n)
    modes = [(i, j, k) for i in range(n) for j in range(n) for k in range(n)]
    kvec = {m: np.array([sgn[m[0]], sgn[m[1]], sgn[m[2]]], float) * dk for m in modes}
    W = {m: po.assignment_function(p, kvec[m]) for m in modes}
    dkb = kb[1] - kb[0]

    def shell(kc):
        lo = kc - dkb * 0.5 if kc > dkb * 0.5 else 0.0
        hi = kc + dkb * 0.5
        return [m for m in modes if lo < np.sqrt(np.sum(kvec[m] ** 2)) <= hi]

    V, Ng = L ** 3, n ** 3
    s1 = shell(kb[1])
    for ik in range(len(kb)):
        s2 = shell(kb[ik])
        tot = 0.0
        for q1 in s1:
            a1 = d[q1] / W[q1]
            for q2 in s2:
                q3 = tuple((-(q1[a] + q2[a])) % n for a in range(3))
                tot += a1 * (d[q2] / W[q2]) * (d[q3] / W[q3])
        expect = (V / Ng) * Ng / (len(s1) * len(s2) * V) * tot
        assert bk_save[ik] == pytest.approx(expect, rel=1e-10, abs=1e-12 * abs(expect) + 1e-300)


def test_box_bk_shotnoise_ijk_numpy():
    """sn_save of bispec.hpp:939-1056 for B_000 from numpy.fft + scipy j0."""
    n, L, N = 12, 600.0, 2000
    p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.05, n_kbin=4, ith_kbin=2)
    kb = po.kbin(p)
    pos = clustered_catalogue(N, L, seed=9, nclump=5, sigma=50.0)
    out, bk_save, sn_save = po.box_bk(p, pos)
    d = po.fourier_transform(p, po.box_density_fluctuation(p, pos))
    nn = d.copy()
    nn[0, 0, 0] += N
    pk, _ = po.calc_power_spectrum(p, d, nn, kb, float(N), 0, 0)
    dk = 2 * np.pi / L
    idx = np.arange(n)
    sgn = np.where(idx < n // 2, idx, idx - n)
    g = np.zeros((n, n, n), complex)
    for i in range(n):
        for j in range(n):
            for k in range(n):
                kv = np.array([sgn[i], sgn[j], sgn[k]], float) * dk
                S = po.shotnoise_function(p, kv)
                W = po.assignment_function(p, kv)
                g[i, j, k] = (d[i, j, k] * np.conj(nn[i, j, k]) - N * S) / W ** 2 / L ** 3
    xi = np.fft.ifftn(g) * g.size
    dr = L / n
    rx, ry, rz = np.meshgrid(sgn * dr, sgn * dr, sgn * dr, indexing="ij")
    r = np.sqrt(rx ** 2 + ry ** 2 + rz ** 2)
    for ik in range(len(kb)):
        ijk = np.sum(spherical_jn(0, kb[2] * r) * spherical_jn(0, kb[ik] * r) * xi) * (L ** 3 / n ** 3)
        expect = N + pk[2] + pk[ik] + ijk
        assert sn_save[ik] == pytest.approx(expect, rel=1e-9, abs=1e-9 * abs(expect))


@pytest.mark.parametrize("ell", [0, 2])
def test_xi_plane_wave_closed_form(ell):
    """calcCorrelationFunction (field.hpp:854-977) of a two-mode field: the
    inverse transform is (|A|^2/V)(e^{i k0.r}/W(k0)^2 + e^{-i k0.r}/W(-k0)^2)
    in closed form; binned with the 0.5 Mpc/h fine rule and P_l(r_z/|r|)."""
    from scipy.special import eval_legendre

    n, L = 16, 400.0
    p = po.param(n, L, po.TSC, rmin=30.0, rmax=150.0, n_rbin=7)
    rb = po.rbin(p)
    f = np.zeros((n, n, n), complex)
    A = 2.0 - 1.5j
    f[2, 1, 3] = A
    f[-2, -1, -3] = np.conj(A)
    xi, nm = po.calc_correlation_function(p, f, f, rb, 0.0, ell, 0)
    dk = 2 * np.pi / L
    k0 = np.array([2, 1, 3]) * dk
    W1 = po.assignment_function(p, k0)
    W2 = po.assignment_function(p, -k0)
    idx = np.arange(n)
    sgn = np.where(idx < n // 2, idx, idx - n) * (L / n)
    rx, ry, rz = np.meshgrid(sgn, sgn, sgn, indexing="ij")
    r = np.sqrt(rx ** 2 + ry ** 2 + rz ** 2)
    phase = k0[0] * rx + k0[1] * ry + k0[2] * rz
    field = abs(A) ** 2 / L ** 3 * (np.exp(1j * phase) / W1 ** 2 + np.exp(-1j * phase) / W2 ** 2)
    with np.errstate(invalid="ignore", divide="ignore"):
        y = np.where(r > 0, eval_legendre(ell, np.where(r > 0, rz / r, 0.0)), 0.0) if ell else np.ones_like(r)
    centre = (r / 0.5 + 0.5).astype(int) * 0.5
    drb = rb[1] - rb[0]
    for j, c in enumerate(rb):
        lo = c - drb * 0.5 if c > drb * 0.5 else 0.0
        sel = (centre > lo) & (centre <= c + drb * 0.5)
        assert nm[j] == sel.sum()
        expect = (field[sel] * y[sel]).sum() / sel.sum() if sel.sum() else 0.0
        assert abs(xi[j] - expect) <= 1e-12 * abs(A) ** 2 / L ** 3 * 2 / min(W1, W2) ** 2


@pytest.mark.parametrize("ells", [(1, 1, 0), (2, 0, 2), (1, 1, 2)])
def test_box_bk_general_ell_bruteforce(ells):
    """bk_save of bispec.hpp:1100-1182 for (l1, l2, L): the direct triangle sum
    sum_m w_m sum_{q1 in S1, q2 in S2} Y(l1, m)(q1) Y(l2, -m)(q2) d(q1) d(q2) d(-q1-q2) / W^3."""
    l1, l2, L_ = ells
    n, L = 10, 600.0
    p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.05, n_kbin=4, ith_kbin=1, ell1=l1, ell2=l2, ELL=L_)
    kb = po.kbin(p)
    pos = clustered_catalogue(2500, L, seed=17, nclump=6, sigma=40.0)
    out, bk_save, sn_save = po.box_bk(p, pos)
    d = po.fourier_transform(p, po.box_density_fluctuation(p, pos))
    dk = 2 * np.pi / L
    idx = np.arange(n)
    sgn = np.where(idx < n // 2, idx, idx - n)
    modes = [(i, j, k) for i in range(n) for j in range(n) for k in range(n)]
    kvec = {m: np.array([sgn[m[0]], sgn[m[1]], sgn[m[2]]], float) * dk for m in modes}
    a = {m: d[m] / po.assignment_function(p, kvec[m]) for m in modes}
    dkb = kb[1] - kb[0]

    def shell(kc):
        lo = kc - dkb * 0.5 if kc > dkb * 0.5 else 0.0
        return [m for m in modes if lo < np.sqrt(np.sum(kvec[m] ** 2)) <= kc + dkb * 0.5]

    def ylm(l, m, v):
        out = np.zeros(2)
        po.lib().or_ylm(l, m, po.dp(np.ascontiguousarray(v, dtype=np.float64)), po.dp(out))
        return out[0] + 1j * out[1]

    weights = []
    for m1 in range(-l1, l1 + 1):
        w = po.wigner3j(l1, l2, L_, 0, 0, 0) * po.wigner3j(l1, l2, L_, m1, -m1, 0)
        w *= (2 * L_ + 1) * (2 * l1 + 1) * (2 * l2 + 1)
        if abs(w) >= 1e-10 and abs(m1) <= l2:
            weights.append((m1, w))
    s1 = shell(kb[1])
    for ik in range(len(kb)):
        s2 = shell(kb[ik])
        tot = 0.0
        for m1, w in weights:
            y1 = {q: ylm(l1, m1, kvec[q]) for q in s1}
            y2 = {q: ylm(l2, -m1, kvec[q]) for q in s2}
            for q1 in s1:
                for q2 in s2:
                    q3 = tuple((-(q1[c] + q2[c])) % n for c in range(3))
                    tot += w * y1[q1] * y2[q2] * a[q1] * a[q2] * a[q3]
        expect = tot / (len(s1) * len(s2) * L ** 3) * (L ** 3 / n ** 3) * n ** 3
        assert bk_save[ik] == pytest.approx(expect, rel=1e-9, abs=1e-10 * abs(expect) + 1e-300)


@pytest.mark.parametrize("ells", [(1, 1, 0), (2, 0, 2), (0, 2, 2)])
def test_box_bk_general_ell_shotnoise(ells):
    """sn_save of bispec.hpp:912-1073 for (l1, l2, L) from numpy.fft, scipy j_l and the
    pinned Y_lm: the P terms enter only with l2 = 0 (row bin) or l1 = 0 (column bin)."""
    from scipy.special import spherical_jn

    l1, l2, L_ = ells
    n, L, N = 10, 600.0, 2000
    p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.05, n_kbin=4, ith_kbin=2, ell1=l1, ell2=l2, ELL=L_)
    kb = po.kbin(p)
    pos = clustered_catalogue(N, L, seed=19, nclump=5, sigma=50.0)
    _, _, sn_save = po.box_bk(p, pos)
    d = po.fourier_transform(p, po.box_density_fluctuation(p, pos))
    nn = d.copy()
    nn[0, 0, 0] += N
    dk = 2 * np.pi / L
    idx = np.arange(n)
    sgn = np.where(idx < n // 2, idx, idx - n)
    g = np.zeros((n, n, n), complex)
    for i in range(n):
        for j in range(n):
            for k in range(n):
                kv = np.array([sgn[i], sgn[j], sgn[k]], float) * dk
                g[i, j, k] = (d[i, j, k] * np.conj(nn[i, j, k]) - N * po.shotnoise_function(p, kv)) \
                    / po.assignment_function(p, kv) ** 2 / L ** 3
    xi = np.fft.ifftn(g) * g.size
    dr = L / n
    R = np.stack(np.meshgrid(sgn * dr, sgn * dr, sgn * dr, indexing="ij"), -1).reshape(-1, 3)
    r = np.linalg.norm(R, axis=1)

    def ylm_grid(l, m):
        out = np.zeros(len(R), complex)
        for q, v in enumerate(R):
            o = np.zeros(2)
            po.lib().or_ylm(l, m, po.dp(np.ascontiguousarray(v)), po.dp(o))
            out[q] = o[0] + 1j * o[1]
        return out

    expect = np.zeros(len(kb), complex)
    for m1 in range(-l1, l1 + 1):
        m2 = -m1
        if abs(m2) > l2:
            continue
        w = po.wigner3j(l1, l2, L_, 0, 0, 0) * po.wigner3j(l1, l2, L_, m1, m2, 0)
        w *= (2 * L_ + 1) * (2 * l1 + 1) * (2 * l2 + 1)
        if abs(w) < 1e-10:
            continue
        if l2 == 0:
            pk, _ = po.calc_power_spectrum(p, d, nn, kb, float(N), l1, m1)
            expect += w * pk[2]
        if l1 == 0:
            pk, _ = po.calc_power_spectrum(p, d, nn, kb, float(N), l2, m2)
            expect += w * pk
        Y = ylm_grid(l1, m1) * ylm_grid(l2, m2)
        for ik in range(len(kb)):
            s = np.sum(spherical_jn(l1, kb[2] * r) * spherical_jn(l2, kb[ik] * r) * xi.ravel() * Y)
            expect[ik] += w * (1j ** (l1 + l2)) * s * (L ** 3 / n ** 3)
    scale = np.abs(expect).max()
    assert np.abs(sn_save - expect).max() <= 1e-8 * scale
