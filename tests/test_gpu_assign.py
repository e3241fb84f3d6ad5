"""HIP mass assignment (hm_assign, hm_cell_coords) against the oracle's
calcField restatement (field.hpp:47-184): stencil cell indices bit for bit,
grid values to the fixed-point resolution, bitwise determinism."""
import numpy as np
import pytest

import hitomi_amd as hm
from oracle import pyoracle as po

from .conftest import clustered_catalogue, uniform_catalogue

pytestmark = pytest.mark.gpu

SCHEMES = [("NGP", po.NGP), ("CIC", po.CIC), ("TSC", po.TSC)]


def edge_catalogue(n, L):
    """Positions on cell centres, half-cell edges, the box faces and just below L."""
    dx = L / n
    vals = np.array([0.0, 1e-12, 0.5 * dx, dx, 1.5 * dx, (n / 2) * dx, (n - 1.5) * dx, (n - 0.5) * dx, L - dx * 1e-9,
                     np.nextafter(L, 0.0), 3.999999999 * dx, 4.5 * dx])
    g = np.array(np.meshgrid(vals, vals, vals, indexing="ij")).reshape(3, -1).T
    return np.ascontiguousarray(g)


@pytest.mark.parametrize("name,code", SCHEMES)
def test_cell_coords_bit_exact(ctx, name, code):
    for n, L in [(32, 1000.0), (64, 2000.0), (128, 1000.0)]:
        pos = np.concatenate([edge_catalogue(n, L), uniform_catalogue(20000, L, seed=n, rsd_sigma=5.0)])
        got = ctx.cell_coords(hm.make_mesh(n, L, name), pos)
        ref = po.cell_coords(po.param(n, L, code), pos)
        assert np.array_equal(got, ref), (name, n)


@pytest.mark.parametrize("name,code", SCHEMES)
def test_assign_matches_oracle(ctx, name, code):
    n, L = 32, 500.0
    pos = np.concatenate([edge_catalogue(n, L), clustered_catalogue(30000, L, seed=3)])
    re, im = ctx.assign(hm.make_mesh(n, L, name), pos)
    ref = po.calc_field(po.param(n, L, code), pos)
    factor = (n / L) ** 3
    # fixed point: each of the <= 27 N contributions is rounded to 2^-(62-23) here
    atol = 1e-12 * np.abs(ref.real).max() + 1e-9 * factor
    assert np.abs(re - ref.real).max() <= atol
    assert np.all(im == 0)


def test_assign_complex_weights(ctx):
    n, L = 16, 300.0
    rng = np.random.default_rng(8)
    pos = uniform_catalogue(5000, L, seed=9)
    w = rng.standard_normal((5000, 2))
    re, im = ctx.assign(hm.make_mesh(n, L, "CIC"), pos, w, wabs_sum=np.abs(w).sum())
    ref = po.calc_field(po.param(n, L, po.CIC), pos, w)
    scale = np.abs(ref).max()
    assert np.abs(re - ref.real).max() < 1e-10 * scale
    assert np.abs(im - ref.imag).max() < 1e-10 * scale


def test_assign_order_independent_bitwise(ctx):
    n, L = 64, 1000.0
    pos = clustered_catalogue(200000, L, seed=21)
    perm = np.random.default_rng(1).permutation(pos.shape[0])
    a, _ = ctx.assign(hm.make_mesh(n, L, "TSC"), pos)
    b, _ = ctx.assign(hm.make_mesh(n, L, "TSC"), pos[perm])
    c, _ = ctx.assign(hm.make_mesh(n, L, "TSC"), pos)
    assert np.array_equal(a, b) and np.array_equal(a, c)


@pytest.mark.parametrize("name", ["NGP", "CIC", "TSC"])
def test_sorted_gather_equals_atomic_scatter_bitwise(ctx, name):
    """The sort-by-cell gather and the atomic scatter round every stencil
    contribution to the same fixed-point value: identical grids, including the
    spilled / dropped border contributions (edge catalogue)."""
    for n, L in [(16, 100.0), (64, 1000.0)]:
        pos = np.concatenate([edge_catalogue(n, L), clustered_catalogue(50000, L, seed=n)])
        a, _ = ctx.assign(hm.make_mesh(n, L, name), pos)
        b, _ = ctx.assign(hm.make_mesh(n, L, name), pos, atomic=True)
        assert np.array_equal(a, b), (name, n)
    w = np.random.default_rng(3).standard_normal((pos.shape[0], 2))
    a = ctx.assign(hm.make_mesh(n, L, name), pos, w, wabs_sum=np.abs(w).sum())
    b = ctx.assign(hm.make_mesh(n, L, name), pos, w, wabs_sum=np.abs(w).sum(), atomic=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_assign_empty_catalogue(ctx):
    re, im = ctx.assign(hm.make_mesh(16, 100.0, "TSC"), np.zeros((0, 3)))
    assert np.all(re == 0)


@pytest.mark.parametrize("nbytes", [16, 4096 + 8, 3 * 2**20 + 5])
def test_device_copy(ctx, nbytes):
    rng = np.random.default_rng(nbytes)
    src = rng.integers(0, 256, nbytes, dtype=np.uint8)
    a = ctx.array(src)
    b = ctx.empty((nbytes,), np.uint8)
    b.copy_from(a)
    assert np.array_equal(b.download(), src)


def test_sort_reuse_with_border_particles_matches_oracle(ctx):
    """A second assignment of the same positions moves only its weights into cell order (each particle's place
    is kept from the first sort).  The reconstructed-box bispectrum's shot-noise density D + alpha^2 R reuses
    the sort of the density D - alpha R; with an edge catalogue (particles whose stencil spills into the next
    row or leaves the mesh) its sums equal the oracle's, which assigns both fields afresh, to FP64 rounding."""
    from oracle import pyoracle as po

    n, L = 32, 400.0
    posD = np.concatenate([edge_catalogue(n, L), clustered_catalogue(20000, L, seed=41)])
    posR = np.concatenate([edge_catalogue(n, L)[::3], uniform_catalogue(40000, L, seed=42)])
    kw = dict(n_kbin=4, kmin=0.03, kmax=0.12)
    alpha = posD.shape[0] / posR.shape[0]
    plan = hm.BoxPlan(ctx, n, L, "TSC", pk_ells=(0,), bk=True, precision=hm.FP64, **kw)
    res = plan.execute_recon(posD, posR, alpha)
    plan.close()
    for ith in range(4):
        p = po.param(n, L, po.TSC, ith_kbin=ith, **kw)
        _, bk, sn = po.box_recon_bk(p, posD, posR, alpha)
        got_sn = res["sn_sum"][ith, :, 0] + 1j * res["sn_sum"][ith, :, 1]
        got_bk = res["bk_sum"][ith, :, 0] + 1j * res["bk_sum"][ith, :, 1]
        assert np.abs(got_sn - sn).max() <= 1e-10 * np.abs(sn).max()
        assert np.abs(got_bk - bk).max() <= 1e-10 * np.abs(bk).max()
