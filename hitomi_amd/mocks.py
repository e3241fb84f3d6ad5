"""Seeded Zel'dovich mock catalogues for periodic boxes (host side, numpy).

The reference has no mock generator (its only synthetic input is an unseeded
uniform random catalogue, struct/measurement/nbody/pk/calc_random.py:38-49);
the measurements here need reproducible clustered catalogues of a stated
shape.  A Gaussian field with power spectrum P(k) is drawn on an n_side^3
lattice from a seeded normal white-noise field, and every lattice particle is
moved by the Zel'dovich displacement Psi(k) = i k / k^2 delta(k), optionally
with the linear redshift-space term f (Psi . z) z along the z axis, then
wrapped into [0, L) once per axis as calcPeriodicBoundary does
(particleGadget.hpp:565-591).
"""
from __future__ import annotations

import numpy as np


def pk_powerlaw(k, amp=2.0e4, k0=0.05, slope=-1.5, kcut=0.5):
    """P(k) = amp (k/k0)^slope exp(-(k/kcut)^2), P(0) = 0 (k in h/Mpc)."""
    k = np.asarray(k, dtype=np.float64)
    out = np.zeros_like(k)
    m = k > 0
    out[m] = amp * (k[m] / k0) ** slope * np.exp(-(k[m] / kcut) ** 2)
    return out


def pk_bao(k, wiggle=0.05, period=0.0628):
    """The power law above times a BAO-like wiggle 1 + wiggle sin(2 pi k / period)."""
    k = np.asarray(k, dtype=np.float64)
    return pk_powerlaw(k) * (1.0 + wiggle * np.sin(2.0 * np.pi * k / period))


def gaussian_field_k(n_side, L, seed, pk=pk_powerlaw):
    """Half-spectrum delta(k) (numpy rfftn layout) of a Gaussian field with
    <|delta_k|^2> = n^3 P(k) / V_cell, i.e. P = V |delta_k|^2 / n^6."""
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((n_side,) * 3)
    wk = np.fft.rfftn(white)
    kx = 2.0 * np.pi * np.fft.fftfreq(n_side, d=L / n_side)
    kz = 2.0 * np.pi * np.fft.rfftfreq(n_side, d=L / n_side)
    k2 = kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2
    amp = np.sqrt(pk(np.sqrt(k2)) / (L / n_side) ** 3)
    return wk * amp, (kx, kz, k2)


def zeldovich(n_side, L, seed, pk=pk_powerlaw, f_growth=0.0, bias_shift=0.0):
    """(n_side^3, 3) float64 positions in [0, L)^3."""
    dk, (kx, kz, k2) = gaussian_field_k(n_side, L, seed, pk)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = np.where(k2 > 0, 1.0 / k2, 0.0)
    q = (np.arange(n_side) + 0.5) * (L / n_side)
    pos = np.empty((n_side,) * 3 + (3,), dtype=np.float64)
    grids = np.meshgrid(q, q, q, indexing="ij")
    comps = [kx[:, None, None], kx[None, :, None], kz[None, None, :]]
    for a in range(3):
        psi = np.fft.irfftn(1j * comps[a] * inv * dk, s=(n_side,) * 3)
        if a == 2 and f_growth:
            psi = psi * (1.0 + f_growth)
        pos[..., a] = grids[a] + psi + bias_shift
    pos = pos.reshape(-1, 3)
    pos = np.mod(pos, L)
    pos[pos >= L] -= L
    return np.ascontiguousarray(pos)


def poisson_sizes(n_mocks, mean, lo, hi, seed, scatter=0.0):
    """Per-mock particle counts ~ Poisson(lambda) truncated to [lo, hi] by
    redrawing; lambda = mean, or with scatter > 0 lambda ~ Gamma with that
    mean and relative width (a Gamma-Poisson mixture: uneven sizes)."""
    rng = np.random.default_rng(seed)

    def draw(k):
        if scatter > 0.0:
            shape = 1.0 / scatter ** 2
            return rng.poisson(rng.gamma(shape, mean / shape, k))
        return rng.poisson(mean, k)

    out = draw(n_mocks)
    bad = (out < lo) | (out > hi)
    while bad.any():
        out[bad] = draw(int(bad.sum()))
        bad = (out < lo) | (out > hi)
    return out
