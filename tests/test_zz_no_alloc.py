"""hm_assign_noalloc (the mass assignment that allocates nothing: its sort work
lives in the library's fixed device scratch) against hm_assign, called through
the C ABI: bitwise equal grids on two calls in a row, the second with a smaller
catalogue on a smaller mesh (the shared scratch keeps the first call's
contents), with and without complex weights; inputs beyond the scratch are
rejected with an error before anything is launched."""
import ctypes as ct

import numpy as np
import pytest

import hitomi_amd as hm

from .conftest import clustered_catalogue

pytestmark = pytest.mark.gpu


def _both(ctx, mesh, pos, w=None):
    d = ctx.array(pos)
    dw = ctx.array(w) if w is not None else None
    shape = tuple(mesh.n_mesh)
    out = {}
    for name, fn in (("alloc", ctx.lib.hm_assign), ("noalloc", ctx.lib.hm_assign_noalloc)):
        re = ctx.empty(shape, np.float64)
        im = ctx.empty(shape, np.float64) if w is not None else None
        wabs = float(np.abs(w).sum()) if w is not None else 0.0
        rc = fn(ctx.h, ct.byref(mesh), d.ptr, dw.ptr if dw else None, pos.shape[0], wabs, re.ptr,
                im.ptr if im else None)
        ctx.check(rc, name)
        ctx.sync()
        out[name] = (re.download(), im.download() if im is not None else None)
        re.free()
        if im is not None:
            im.free()
    d.free()
    if dw is not None:
        dw.free()
    return out


@pytest.mark.parametrize("scheme", ["NGP", "CIC", "TSC"])
def test_noalloc_equals_assign_twice(ctx, scheme):
    first = clustered_catalogue(600_000, 1000.0, seed=11)
    second = clustered_catalogue(50_000, 500.0, seed=12)
    for n, L, pos in ((128, 1000.0, first), (64, 500.0, second)):
        out = _both(ctx, hm.make_mesh(n, L, scheme), pos)
        assert np.array_equal(out["alloc"][0], out["noalloc"][0]), (scheme, n)
        assert np.abs(out["alloc"][0]).sum() > 0


def test_noalloc_complex_weights_twice(ctx):
    rng = np.random.default_rng(4)
    for n, L, npart in ((96, 800.0, 300_000), (32, 400.0, 20_000)):
        pos = clustered_catalogue(npart, L, seed=npart)
        w = rng.standard_normal((npart, 2))
        out = _both(ctx, hm.make_mesh(n, L, "TSC"), pos, w)
        assert np.array_equal(out["alloc"][0], out["noalloc"][0])
        assert np.array_equal(out["alloc"][1], out["noalloc"][1])


def test_noalloc_rejects_oversize(ctx):
    mesh = hm.make_mesh(64, 500.0, "TSC")
    npart = 5_000_000  # 32 B each without weights: 160 MB, more than the 128 MiB scratch
    out = ctx.empty((64, 64, 64), np.float64)
    pos = ctx.array(np.zeros((npart, 3)))  # a valid catalogue: the size check rejects it before any launch
    rc = ctx.lib.hm_assign_noalloc(ctx.h, ct.byref(mesh), pos.ptr, None, npart, 0.0, out.ptr, None)
    assert rc != 0
    assert b"scratch" in ctx.lib.hm_last_error(ctx.h)
    out.free()
    pos.free()
