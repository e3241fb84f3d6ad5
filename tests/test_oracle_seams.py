"""Pins the oracle's single-seam functions (the pieces its drivers compose,
exported for the library's per-seam entry points) by closed forms, by scipy
and by relations to already-pinned oracle functions:

  - inverse transform [field.hpp:328-344]: the forward transform followed by
    the inverse one is the identity (dV N / V = 1);
  - triple sum [bispec.hpp:1159-1168]: plane waves e^{i k_a x} e^{i k_b x}
    e^{i k_c x} sum to V when k_a + k_b + k_c = 0 and to 0 otherwise;
  - shot-noise Bessel sweep [bispec.hpp:1036-1056]: scipy's j_l with the
    analytic Y_red(1, 0)(r) = z / |r| and the i^(l1+l2) = -1 sign;
  - Bessel inverse transform [field.hpp:403-445]: one Fourier mode a at k0
    comes back as j_0(|k0| r) a / W(k0) / V e^{i k0 x} (scipy's j_0);
  - 3PCF correlation function [field.hpp:979-1113]: with uniform bins its
    edges equal calcCorrelationFunction's, so xi_3pcf = xi / (n_mode dV)
    against the plane-wave-pinned calcCorrelationFunction;
  - particle sums [field.hpp:1164-1240]: Y_red(2, 0)(z) = 1, Y_red(2, 0)(x) =
    -1/2, Y_red(1, 1)(x) = -1/sqrt(2);
  - survey volume [field.hpp:449-503]: one particle at each cell centre of
    a box (NGP) gives V exactly, of half the box V / 2.
"""
import numpy as np
from scipy.special import spherical_jn

from oracle import pyoracle as po


def _grid(p, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=po.shape(p)) + 1j * rng.normal(size=po.shape(p))


def _rvec(p):
    out = []
    for a in range(3):
        n, L = p.n_mesh[a], p.boxsize[a]
        d = L / n
        i = np.arange(n)
        out.append(np.where(i < n // 2, i * d, (i - n) * d))
    return np.meshgrid(*out, indexing="ij")


def _kvec(p):
    out = []
    for a in range(3):
        n, L = p.n_mesh[a], p.boxsize[a]
        d = 2 * np.pi / L
        i = np.arange(n)
        out.append(np.where(i < n // 2, i * d, (i - n) * d))
    return np.meshgrid(*out, indexing="ij")


def test_inverse_transform_undoes_forward():
    p = po.param((8, 10, 12), (100.0, 130.0, 170.0))
    f = _grid(p, 1)
    back = po.inverse_fourier_transform(p, po.fourier_transform(p, f))
    assert np.abs(back - f).max() < 1e-12 * np.abs(f).max()


def test_triple_sum_plane_waves():
    p = po.param((8, 10, 12), (100.0, 130.0, 170.0))
    V = 100.0 * 130.0 * 170.0
    L = np.array([100.0, 130.0, 170.0])
    x = np.stack(np.meshgrid(*[np.arange(n) * L[a] / n for a, n in enumerate((8, 10, 12))], indexing="ij"), -1)

    def wave(nv):
        return np.exp(1j * (x @ (2 * np.pi * np.asarray(nv) / L)))

    a, b = (1, -2, 3), (2, 1, -1)
    c = tuple(-(u + v) for u, v in zip(a, b))
    s = po.triple_sum(p, wave(a), wave(b), wave(c))
    assert abs(s - V) < 1e-10 * V
    s0 = po.triple_sum(p, wave(a), wave(b), wave((0, 0, 1)))
    assert abs(s0) < 1e-10 * V


def test_sn_bessel_sweep_scipy():
    p = po.param((12, 12, 16), (200.0, 200.0, 260.0))
    rx, ry, rz = _rvec(p)
    r = np.sqrt(rx * rx + ry * ry + rz * rz)
    xi = _grid(p, 2)
    dV = 200.0 * 200.0 * 260.0 / (12 * 12 * 16)
    k1, k2 = 0.05, 0.08
    # l1 = l2 = 0
    ref0 = dV * np.sum(spherical_jn(0, k1 * r) * spherical_jn(0, k2 * r) * xi)
    got0 = po.sn_bessel_sweep(p, xi, 0, 0, 0, 0, k1, k2)
    assert abs(got0 - ref0) < 1e-8 * np.abs(ref0)
    # l1 = l2 = 1, m = 0: Y_red(1, 0)(r) = z / |r| (0 at r = 0), i^2 = -1
    mu = np.where(r > 0, rz / np.where(r > 0, r, 1.0), 0.0)
    ref1 = -dV * np.sum(spherical_jn(1, k1 * r) * spherical_jn(1, k2 * r) * xi * mu * mu)
    got1 = po.sn_bessel_sweep(p, xi, 1, 0, 1, 0, k1, k2)
    assert abs(got1 - ref1) < 1e-8 * np.abs(ref1)


def test_bessel_ifft_single_mode():
    p = po.param((8, 8, 10), (150.0, 150.0, 200.0), assign=po.CIC)
    idx = (1, 7, 2)  # kx > 0, ky < 0 (index 7 of 8), kz > 0
    d = np.zeros(po.shape(p), dtype=np.complex128)
    a = 0.7 - 0.4j
    d[idx] = a
    kx, ky, kz = _kvec(p)
    k0 = np.array([kx[idx], ky[idx], kz[idx]])
    W = po.assignment_function(p, k0)
    rmag = 37.0
    out = po.ifft_for_3pcf(p, d, rmag, 0, 0, 0)
    L = np.array([150.0, 150.0, 200.0])
    x = np.stack(np.meshgrid(*[np.arange(n) * L[q] / n for q, n in enumerate((8, 8, 10))], indexing="ij"), -1)
    ref = spherical_jn(0, np.linalg.norm(k0) * rmag) * a / W / np.prod(L) * np.exp(1j * (x @ k0))
    assert np.abs(out - ref).max() < 1e-9 * np.abs(ref).max()


def test_corr_for_3pcf_equals_2pt_scaled():
    p = po.param((12, 12, 12), (240.0, 240.0, 240.0), assign=po.TSC, n_rbin=6)
    f = _grid(p, 3)
    rb = np.linspace(20.0, 70.0, 6)
    xi2, n2 = po.calc_correlation_function(p, f, f, rb, 5.0, 0, 0)
    xi3, n3 = po.corr_for_3pcf(p, f, f, rb, 5.0, 0, 0, 0, 0)
    assert np.array_equal(n2, n3)
    dV = 240.0 ** 3 / 12 ** 3
    ok = n2 > 0
    assert np.abs(xi3[ok] - xi2[ok] / (n2[ok] * dV)).max() < 1e-12 * np.abs(xi3[ok]).max()


def test_particle_sum_closed_forms():
    rng = np.random.default_rng(4)
    w = rng.random(50) + 0.5
    z = np.tile([0.0, 0.0, 2.5], (50, 1))
    x = np.tile([3.0, 0.0, 0.0], (50, 1))
    assert abs(po.particle_sum(z, w, 2, 0, 2) - np.sum(w ** 2)) < 1e-12 * np.sum(w ** 2)
    assert abs(po.particle_sum(x, w, 2, 0, 2) + 0.5 * np.sum(w ** 2)) < 1e-12 * np.sum(w ** 2)
    s = po.particle_sum(x, w, 1, 1, 3)
    assert abs(s + np.sum(w ** 3) / np.sqrt(2.0)) < 1e-12 * np.sum(w ** 3)
    assert abs(po.particle_sum(x, None, 0, 0, 2) - 50.0) < 1e-12


def test_survey_volume_closed_forms():
    n, L = (6, 8, 10), (60.0, 80.0, 100.0)
    p = po.param(n, L, assign=po.NGP)
    d = np.array(L) / np.array(n)
    c = np.stack(np.meshgrid(*[np.arange(m) * d[a] for a, m in enumerate(n)], indexing="ij"), -1).reshape(-1, 3)
    V = np.prod(L)
    assert abs(po.survey_volume(p, c) - V) < 1e-12 * V
    assert abs(po.survey_volume(p, c, np.full(c.shape[0], 2.0)) - V) < 1e-12 * V
    half = c[c[:, 0] < L[0] / 2]
    assert abs(po.survey_volume(p, half) - V / 2) < 1e-12 * V
