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


# ---------------------------------------------------------------------------
# Survey-geometry mocks: data Zel'dovich-displaced from a seeded Gaussian field
# and kept where they land inside a sky patch, randoms uniform in the patch's
# volume, FKP weights w = 1 / (1 + nbar P0) with the patch's mean density
# (the reference's catalogues come as `x y z w` text from
# struct/data/boss/fits2dat_galaxy_weight.py; the survey measurement reads them
# through particle.hpp:93-154).
# ---------------------------------------------------------------------------
def sky_patch_volume(patch):
    """(RA range in rad) (sin Dec_max - sin Dec_min) (rmax^3 - rmin^3) / 3."""
    deg = np.pi / 180.0
    return ((patch.ra_max - patch.ra_min) * deg * (np.sin(patch.dec_max * deg) - np.sin(patch.dec_min * deg))
            * (patch.rmax ** 3 - patch.rmin ** 3) / 3.0)


def in_sky_patch(pos, patch):
    """Mask of the positions inside the patch (r, RA in [0, 360), Dec from the Cartesian position)."""
    pos = np.asarray(pos, dtype=np.float64)
    r = np.sqrt(np.sum(pos * pos, axis=1))
    with np.errstate(invalid="ignore", divide="ignore"):
        dec = np.degrees(np.arcsin(np.clip(pos[:, 2] / r, -1.0, 1.0)))
    ra = np.mod(np.degrees(np.arctan2(pos[:, 1], pos[:, 0])), 360.0)
    return ((r >= patch.rmin) & (r < patch.rmax) & (ra >= patch.ra_min) & (ra < patch.ra_max)
            & (dec >= patch.dec_min) & (dec < patch.dec_max))


def sky_patch_bounds(patch, samples=181):
    """Cartesian bounding box of the patch from a dense sample of its surface."""
    deg = np.pi / 180.0
    ra = np.linspace(patch.ra_min, patch.ra_max, samples) * deg
    dec = np.linspace(patch.dec_min, patch.dec_max, samples) * deg
    r = np.array([patch.rmin, patch.rmax])
    R, A, D = np.meshgrid(r, ra, dec, indexing="ij")
    xyz = np.stack([R * np.cos(D) * np.cos(A), R * np.cos(D) * np.sin(A), R * np.sin(D)], axis=-1).reshape(-1, 3)
    return xyz.min(axis=0), xyz.max(axis=0)


def survey_catalogue(ctx, patch, n_data, n_rand, seed, P0=1.0e4, n_side=256, margin=30.0, pk=None,
                     randoms_on_device=True):
    """A survey-geometry mock: returns dict(posD [n_data][3] host, wD host, posR, wR (device arrays when
    randoms_on_device, else host), nbar, volume, box (the Zel'dovich box side)).

    Data: a periodic Zel'dovich box of side max(extent) + 2 margin generated on the device
    (hm_mock_zeldovich, seeded site subset of an n_side^3 lattice, P(k) = pk), moved to the patch's
    bounding box, the particles inside the patch kept in the generator's (seeded, random) order, the
    first n_data of them.  Randoms: hm_mock_sky_randoms with seed + 1.  Weights: FKP with the mean
    density nbar = n_data / V_patch."""
    import hitomi_amd as hm

    pk = pk or dict(amp=2.0e4, k0=0.05, slope=-1.5, kcut=0.5)
    vol = sky_patch_volume(patch)
    nbar = n_data / vol
    lo, hi = sky_patch_bounds(patch)
    side = float(np.max(hi - lo)) + 2.0 * margin
    n_box = int(np.ceil(1.3 * nbar * side ** 3))
    if n_box > n_side ** 3:
        raise ValueError(f"survey_catalogue: {n_box} lattice sites needed, n_side^3 = {n_side ** 3}")
    cfg = hm.mock_config(n_side, side, **pk)
    dbox = hm.mock_zeldovich(ctx, cfg, int(seed), n_box)
    ctx.sync()
    box = dbox.download()
    dbox.free()
    pos = box + (lo - margin)
    sel = pos[in_sky_patch(pos, patch)]
    if sel.shape[0] < n_data:
        raise ValueError(f"survey_catalogue: only {sel.shape[0]} particles inside the patch")
    posD = np.ascontiguousarray(sel[:n_data])
    w = 1.0 / (1.0 + nbar * P0)
    wD = np.full(n_data, w)
    if randoms_on_device:
        posR = hm.mock_sky_randoms(ctx, patch, int(seed) + 1, n_rand)
        wR = ctx.array(np.full(n_rand, w))
    else:
        d = hm.mock_sky_randoms(ctx, patch, int(seed) + 1, n_rand)
        ctx.sync()
        posR, wR = d.download(), np.full(n_rand, w)
        d.free()
    return dict(posD=posD, wD=wD, posR=posR, wR=wR, nbar=nbar, volume=vol, box=side)
