"""Covariance stage of the configuration-space analysis, on the device.

The reference's stage (struct/covariance/boss/modified_gravity/
calc_data_and_cov.py) reads every mock's 2PCF multipoles xi_l (l = 0, 2, 4)
and 3PCF multipoles zeta_l1l2L ((0,0,0), (1,1,0), (2,0,2), (1,1,2)) back from
text files, divides them bin by bin by the window monopoles w_0 (xi0_window.dat)
and w_000 (zeta000_%02d_window.dat) (:72-105), writes the mock means with their
standard deviations (:142-220) and numpy.cov of every pair of blocks
(:284-307).  Those blocks are the blocks of one covariance of the concatenated
vector, which is what :class:`Covariance` accumulates: each pushed vector is
divided by the window vector and added to FP64 sums on the device
(hm_cov_*), with no host synchronisation per mock; the result is read back once.

Vectors come either from the plans directly (:meth:`Covariance.push` with a
dict of block arrays, or a device vector) or from the files the drop-in
programs write (:func:`from_measurements`), laid out as the reference lays
them out: a 3PCF block is the n_rbin x n_rbin matrix of rows ith_rbin
(zeta%d%d%d_%02d.dat, column 2) flattened row-major (:91-98).
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

# the multipoles the reference's stage combines (calc_data_and_cov.py:72, 88)
XI_ELLS = (0, 2, 4)
ZETA_MULTIPOLES = ((0, 0, 0), (1, 1, 0), (2, 0, 2), (1, 1, 2))


def xi_name(ell):
    return f"xi{ell}"


def zeta_name(l1, l2, L):
    return f"zeta{l1}{l2}{L}"


class Covariance:
    """N - 1 covariance of window-normalised data vectors made of named blocks.

    blocks: [(name, length), ...] in vector order; norm: {name: window values}
    for the blocks to divide (others are taken as they are); keep_rows: how
    many normalised vectors stay on the device for :meth:`vectors`."""

    def __init__(self, ctx, blocks, norm=None, keep_rows=0):
        self.ctx = ctx
        self.names = [b[0] for b in blocks]
        self.sizes = [int(b[1]) for b in blocks]
        if len(set(self.names)) != len(self.names):
            raise ValueError("block names must be distinct")
        off = np.concatenate([[0], np.cumsum(self.sizes)]).astype(int)
        self.slices = {n: slice(int(off[i]), int(off[i + 1])) for i, n in enumerate(self.names)}
        self.dim = int(off[-1])
        norm = norm or {}
        for k in norm:
            if k not in self.slices:
                raise ValueError(f"window given for an unknown block {k!r}")
        self.norm = None
        if norm:
            w = np.ones(self.dim)
            for k, v in norm.items():
                v = np.asarray(v, dtype=np.float64).ravel()
                if v.size != self.sizes[self.names.index(k)]:
                    raise ValueError(f"window of block {k!r} has {v.size} values, the block {self.sizes[self.names.index(k)]}")
                w[self.slices[k]] = v
            self.norm = w
        self.keep_rows = int(keep_rows)
        nptr = self.norm.ctypes.data if self.norm is not None else None
        self.h = ctx.lib.hm_cov_create(ctx.h, self.dim, nptr, self.keep_rows)
        if not self.h:
            ctx.raise_error("hm_cov_create")
        self._res = None

    def close(self):
        if self.h:
            self.ctx.lib.hm_cov_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def concat(self, parts):
        """The vector of a dict {name: values} (3PCF blocks as n x n or flat)."""
        v = np.empty(self.dim)
        for n in self.names:
            a = np.asarray(parts[n], dtype=np.float64).ravel()
            if a.size != self.sizes[self.names.index(n)]:
                raise ValueError(f"block {n!r}: {a.size} values, expected {self.sizes[self.names.index(n)]}")
            v[self.slices[n]] = a
        return v

    def push(self, vec):
        """One mock: a dict of blocks, a host vector, or a device vector (DeviceArray of dim FP64)."""
        self._res = None
        if isinstance(vec, dict):
            vec = self.concat(vec)
        if hasattr(vec, "ptr"):
            if int(np.prod(vec.shape)) != self.dim:
                raise ValueError("device vector of the wrong length")
            self.ctx.check(self.ctx.lib.hm_cov_push(self.ctx.h, self.h, vec.ptr), "hm_cov_push")
            return
        v = np.ascontiguousarray(vec, dtype=np.float64).ravel()
        if v.size != self.dim:
            raise ValueError(f"vector of length {v.size}, expected {self.dim}")
        self.ctx.check(self.ctx.lib.hm_cov_push_host(self.ctx.h, self.h, v.ctypes.data), "hm_cov_push_host")

    @property
    def count(self):
        return self.ctx.lib.hm_cov_count(self.h)

    def vectors(self):
        rows = min(self.count, self.keep_rows)
        out = np.zeros((max(rows, 0), self.dim))
        if rows > 0:
            self.ctx.check(self.ctx.lib.hm_cov_vectors(self.ctx.h, self.h, out.ctypes.data, rows), "hm_cov_vectors")
        return out

    def result(self):
        """(mean [dim], covariance [dim][dim]), the N - 1 estimator of numpy.cov."""
        if self._res is None:
            mean = np.zeros(self.dim)
            cov = np.zeros((self.dim, self.dim))
            self.ctx.check(self.ctx.lib.hm_cov_result(self.ctx.h, self.h, mean.ctypes.data, cov.ctypes.data),
                           "hm_cov_result")
            self._res = (mean, cov)
        return self._res

    def reset(self):
        self._res = None
        self.ctx.check(self.ctx.lib.hm_cov_reset(self.ctx.h, self.h), "hm_cov_reset")

    def block(self, a, b):
        """np.cov(X_a, X_b)[0:n_a, n_a:n_a + n_b] of the reference (calc_data_and_cov.py:288-307)."""
        return self.result()[1][self.slices[a], self.slices[b]]

    def mean(self, name):
        return self.result()[0][self.slices[name]]

    def std(self, name):
        """np.std(..., ddof=1) of the block's entries over the mocks."""
        c = self.result()[1]
        s = self.slices[name]
        return np.sqrt(np.diag(c)[s])

    def write(self, outdir, rbin_2pcf=None, rbin_3pcf=None):
        """The reference's output files (calc_data_and_cov.py:142-148, 199-206, 284-307):
        xi%d_mean.dat, zeta%d%d%d_mean.dat and cov_<a>_<b>.dat for every block pair."""
        os.makedirs(outdir, exist_ok=True)
        for n in self.names:
            if n.startswith("xi") and rbin_2pcf is not None:
                np.savetxt(os.path.join(outdir, f"{n}_mean.dat"),
                           np.array([rbin_2pcf, self.mean(n), self.std(n)]).T, fmt="%.5f \t %.7e \t %.7e")
            elif n.startswith("zeta") and rbin_3pcf is not None:
                ri, rj = zeta_rbins(rbin_3pcf)
                np.savetxt(os.path.join(outdir, f"{n}_mean.dat"),
                           np.array([ri, rj, self.mean(n), self.std(n)]).T, fmt="%.5f \t %.5f \t %.7e \t %.7e")
        for a in self.names:
            for b in self.names:
                np.savetxt(os.path.join(outdir, f"cov_{a}_{b}.dat"), self.block(a, b))


def zeta_rbins(rbin_3pcf):
    """(r_i, r_j) of the flattened 3PCF entries, from the n x n matrix of the
    second column of zeta000_%02d.dat (calc_data_and_cov.py:117-123)."""
    r = np.asarray(rbin_3pcf, dtype=np.float64)
    n = r.shape[0]
    return (r.T).reshape(n * n), r.reshape(n * n)


# ---------------------------------------------------------------------------
# the reference's file layout (what hitomi_survey writes per realisation)
# ---------------------------------------------------------------------------

def read_xi(d, ell, recon=False):
    """Column 1 of xi%d[_recon].dat."""
    return np.loadtxt(os.path.join(d, f"xi{ell}{'_recon' if recon else ''}.dat"), usecols=(1,), unpack=True)


def read_xi_window(d, recon=False):
    """Column 1 of xi0_window.dat or xi0_recon_window.dat (calc_data_and_cov.py:66-69)."""
    return np.loadtxt(os.path.join(d, f"xi0{'_recon' if recon else ''}_window.dat"), usecols=(1,), unpack=True)


def read_zeta(d, l1, l2, L, nr, window=False):
    """Column 2 of zeta%d%d%d_%02d[_window].dat for ith_rbin = 0 .. nr - 1, as the nr x nr matrix."""
    suf = "_window" if window else ""
    return np.array([np.loadtxt(os.path.join(d, f"zeta{l1}{l2}{L}_{i:02d}{suf}.dat"), usecols=(2,), unpack=True)
                     for i in range(nr)])


def read_rbins(d, nr3, recon=False):
    """(r of xi0.dat, the nr3 x nr3 matrix of column 1 of zeta000_%02d.dat) (calc_data_and_cov.py:107-123)."""
    r2 = np.loadtxt(os.path.join(d, f"xi0{'_recon' if recon else ''}.dat"), usecols=(0,), unpack=True)
    r3 = np.array([np.loadtxt(os.path.join(d, f"zeta000_{i:02d}.dat"), usecols=(1,), unpack=True)
                   for i in range(nr3)])
    return r2, r3


def blocks_for(nr2, nr3, xi_ells=XI_ELLS, zeta=ZETA_MULTIPOLES):
    return [(xi_name(l), nr2) for l in xi_ells] + [(zeta_name(*m), nr3 * nr3) for m in zeta]


def windows_for(w0, w000, xi_ells=XI_ELLS, zeta=ZETA_MULTIPOLES):
    """Every xi_l block is divided by w_0, every zeta block by w_000 (calc_data_and_cov.py:78, 95)."""
    w = {xi_name(l): np.asarray(w0, dtype=np.float64) for l in xi_ells}
    w.update({zeta_name(*m): np.asarray(w000, dtype=np.float64).ravel() for m in zeta})
    return w


def from_measurements(ctx, mock_dir, window_2pcf_dir, window_3pcf_dir, n_mocks, nr2, nr3, recon=False,
                      keep_rows=0, first=1):
    """The reference's stage over the files of n_mocks realisations
    mock_dir/%04d (numbered from `first`) and the window measurements:
    returns (Covariance, r of the 2PCF bins, n x n r of the 3PCF bins)."""
    w0 = read_xi_window(window_2pcf_dir, recon)
    w000 = read_zeta(window_3pcf_dir, 0, 0, 0, nr3, window=True)
    cov = Covariance(ctx, blocks_for(nr2, nr3), windows_for(w0, w000), keep_rows=keep_rows)
    for r in range(first, first + n_mocks):
        d = os.path.join(mock_dir, f"{r:04d}")
        parts = {xi_name(l): read_xi(d, l, recon) for l in XI_ELLS}
        parts.update({zeta_name(*m): read_zeta(d, *m, nr3) for m in ZETA_MULTIPOLES})
        cov.push(parts)
    r2, r3 = read_rbins(os.path.join(mock_dir, f"{first:04d}"), nr3, recon)
    return cov, r2, r3
