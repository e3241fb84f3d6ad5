"""Pins the oracle's interlacing (an addition with no counterpart in the
reference, off by default):

  - the interlaced spectrum equals (DFT(a) + exp(i k.h) DFT(b)) / 2 with
    the two grids' DFTs summed explicitly over the cells (8^3), b assigned
    from the positions shifted by h = half a cell and wrapped, for k_z index
    <= n/2 and the conjugate of the mirror mode above (a real field);
  - on a Poisson catalogue it removes the odd-order aliases: in the octant
    with non-negative k (where the window is evaluated without the Q2
    index shift) the mean |d(k)|^2 / (N W^2) above 0.8 k_Nyquist falls from
    the aliased excess to within a few per cent of 1;
  - P_l of the interlaced field equals the binned (|d|^2 - N W^2)/W^2.
"""
import numpy as np

from oracle import pyoracle as po

from .conftest import clustered_catalogue


def test_interlaced_equals_explicit_dft():
    n, L = 8, 400.0
    p = po.param(n, L, po.TSC)
    pos = clustered_catalogue(500, L, seed=3, nclump=3, sigma=30.0)
    h = L / n / 2
    ps = pos + h
    ps[ps >= L] -= L
    a = po.box_density_fluctuation(p, pos) * (L ** 3 / n ** 3)
    b = po.box_density_fluctuation(p, ps) * (L ** 3 / n ** 3)
    idx = np.arange(n)
    sgn = np.where(idx < n // 2, idx, idx - n)
    E = np.exp(-2j * np.pi * np.outer(idx, idx) / n)          # DFT matrix, exp(-i k x)
    dft = lambda g: np.einsum("ai,bj,ck,ijk->abc", E, E, E, g)
    k1 = sgn * 2 * np.pi / L
    KX, KY, KZ = np.meshgrid(k1, k1, k1, indexing="ij")
    expect = 0.5 * (dft(a) + np.exp(1j * (KX + KY + KZ) * h) * dft(b))
    mirror = lambda f: np.conj(np.roll(f[::-1, ::-1, ::-1], 1, axis=(0, 1, 2)))  # f(-k)*
    expect[:, :, n // 2 + 1:] = mirror(expect)[:, :, n // 2 + 1:]
    got = po.box_interlaced_k(p, pos)
    assert np.abs(got - expect).max() <= 1e-12 * np.abs(expect).max()


def test_interlacing_suppresses_odd_aliases_of_poisson_noise():
    n, L, N = 32, 1000.0, 60000
    p = po.param(n, L, po.TSC)
    pos = np.random.default_rng(1).random((N, 3)) * L
    fa = po.fourier_transform(p, po.box_density_fluctuation(p, pos))
    fi = po.box_interlaced_k(p, pos)
    idx = np.arange(n)
    k1 = np.where(idx < n // 2, idx, idx - n) * 2 * np.pi / L
    K = np.stack(np.meshgrid(k1, k1, k1, indexing="ij"), -1)
    W = np.array([po.assignment_function(p, k) for k in K.reshape(-1, 3)]).reshape(n, n, n)
    kmag = np.linalg.norm(K, axis=-1)
    kN = np.pi * n / L
    sel = (kmag > 0.8 * kN) & np.all(K >= 0, axis=-1)
    raw = np.mean(np.abs(fa[sel]) ** 2 / W[sel] ** 2) / N - 1
    inter = np.mean(np.abs(fi[sel]) ** 2 / W[sel] ** 2) / N - 1
    assert raw > 0.15 and abs(inter) < 0.06, (raw, inter)


def test_interlaced_pk_is_binned_with_w2_shotnoise():
    n, L, N = 16, 600.0, 5000
    p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.12, n_kbin=6)
    pos = clustered_catalogue(N, L, seed=5, nclump=5, sigma=30.0)
    got, nm = po.box_interlaced_2pt(p, pos)
    f = po.box_interlaced_k(p, pos)
    kb = po.kbin(p)
    idx = np.arange(n)
    k1 = np.where(idx < n // 2, idx, idx - n) * 2 * np.pi / L
    K = np.stack(np.meshgrid(k1, k1, k1, indexing="ij"), -1).reshape(-1, 3)
    W = np.array([po.assignment_function(p, k) for k in K])
    kmag = np.linalg.norm(K, axis=1)
    v = (np.abs(f.ravel()) ** 2 - N * W ** 2) / W ** 2
    ki = (kmag / 1e-4 + 0.5).astype(int) * 1e-4
    dkb = kb[1] - kb[0]
    for j, kc in enumerate(kb):
        s = (ki > kc - dkb / 2) & (ki <= kc + dkb / 2)
        assert s.sum() == nm[j]
        expect = (L ** 3 / N ** 2) * v[s].mean()
        assert abs(got[j].real - expect) <= 1e-10 * abs(expect)
