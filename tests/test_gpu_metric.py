"""The bispectrum workload at its full size: 512^3 TSC, L = 2000 Mpc/h,
4,000,000 Zel'dovich particles from the seeded Gaussian field with
P(k) = 2e4 (k/0.05)^-1.5 exp(-(k/0.5)^2) [1 + 0.05 sin(2 pi k / 0.0628)]
(a seeded 4,000,000-site subset of a 256^3 lattice, generated on the device),
20 k-bins with centres 0.015 .. 0.205 h/Mpc.

  - the 1/8/27 stencil cell indices of every particle (field.hpp:84-93,
    121-133, 162-176) equal the oracle's bit for bit, for NGP, CIC and TSC;
  - the P_l mode counts (field.hpp:793-837) and the bispectrum shell n_mode
    (field.hpp:361-387) of the FP32 and FP64 plans equal the oracle's count-only
    restatement bit for bit;
  - the FP32 pipeline (re-gridded triple mesh, pruned passes) against the
    plain FP64 one (full 512^3 shell fields, no pruning), whose arithmetic the
    oracle pins at the smaller meshes: P_l within PK_TOL of |P_0| + V/N,
    B_000 within BK_TOL of max|B|;
  - B_110, B_202, zeta_000 and zeta_110 in FP32 against FP64 with the same
    plan shape, within the stated tolerances of the largest |value|.
"""
import numpy as np
import pytest

import hitomi_amd as hm
from oracle import pyoracle as po

from .conftest import report_err

pytestmark = pytest.mark.gpu

N_MESH, BOX, N_SIDE, N_PART, SEED = 512, 2000.0, 256, 4_000_000, 1000
KW = dict(n_kbin=20, kmin=0.015, kmax=0.205)
PK = dict(amp=2.0e4, k0=0.05, slope=-1.5, kcut=0.5, bao_amp=0.05, bao_period=0.0628)

PK_TOL = 1e-6     # FP32 P_l: |dP_l| <= PK_TOL (|P_0| + V/N)
BK_TOL = 5e-6     # FP32 B_000: |dB| <= BK_TOL max|B|
GEN_TOL = 5e-6    # FP32 B_110, B_202: |dB| <= GEN_TOL max|B|
ZETA_TOL = 5e-6   # FP32 zeta: |d norm (zeta - sn)| <= ZETA_TOL max|norm (zeta - sn)|


@pytest.fixture(scope="module")
def catalogue(ctx):
    cfg = hm.mock_config(N_SIDE, BOX, **PK)
    d = hm.mock_zeldovich(ctx, cfg, SEED, N_PART)
    ctx.sync()
    pos = d.download()
    yield d, pos
    d.free()


@pytest.mark.parametrize("name,code", [("NGP", po.NGP), ("CIC", po.CIC), ("TSC", po.TSC)])
def test_metric_cell_coords_bit_exact(ctx, catalogue, name, code):
    _, pos = catalogue
    mesh = hm.make_mesh(N_MESH, BOX, name)
    p = po.param(N_MESH, BOX, code)
    step = 500_000
    for a in range(0, pos.shape[0], step):
        chunk = pos[a:a + step]
        got = ctx.cell_coords(mesh, chunk)
        ref = po.cell_coords(p, chunk)
        assert np.array_equal(got, ref), (name, a)


@pytest.fixture(scope="module")
def b000(ctx, catalogue):
    d, _ = catalogue
    out = {}
    for key, prec, plain in (("fp32", hm.FP32, False), ("fp64_plain", hm.FP64, True)):
        plan = hm.BoxPlan(ctx, N_MESH, BOX, "TSC", pk_ells=(0, 2, 4), bk=True, precision=prec, regrid=not plain,
                          prune=not plain, **KW)
        out[key] = (plan.execute(d), plan.triple_mesh)
        plan.close()
    return out


def test_metric_mode_counts_bit_exact(b000):
    p = po.param(N_MESH, BOX, po.TSC, **KW)
    ref_pk, ref_sh = po.mode_counts(p)
    for key, (res, _) in b000.items():
        assert np.array_equal(res["pk_nmode"], ref_pk), key
        assert np.array_equal(res["shell_nmode"], ref_sh), key


def test_metric_fp32_regrid_vs_plain_fp64(b000):
    r32, m32 = b000["fp32"]
    r64, m64 = b000["fp64_plain"]
    assert m32 == (270, 270, 270) and m64 == (N_MESH,) * 3
    scale = np.abs(r64["pk"][0, :, 0]) + BOX ** 3 / N_PART
    dp = (np.abs(r32["pk"][..., 0] - r64["pk"][..., 0]) / scale[None, :]).max()
    report_err("metric P_l fp32 vs fp64_plain", dp)
    assert dp <= PK_TOL
    b, b64 = r32["bk"][..., 0], r64["bk"][..., 0]
    db = np.abs(b - b64).max() / np.abs(b64).max()
    report_err("metric B_000 fp32 regrid vs fp64 plain", db)
    assert db <= BK_TOL
    assert np.array_equal(b, b.T)


@pytest.mark.parametrize("ells", [(1, 1, 0), (2, 0, 2)])
def test_metric_bk_general_fp32_vs_fp64(ctx, catalogue, ells):
    d, _ = catalogue
    res = {}
    for prec in (hm.FP32, hm.FP64):
        plan = hm.BoxPlan(ctx, N_MESH, BOX, "TSC", pk_ells=(0,), bk=True, bk_ells=ells, precision=prec, **KW)
        res[prec] = plan.execute(d)
        plan.close()
    assert np.array_equal(res[hm.FP32]["shell_nmode"], res[hm.FP64]["shell_nmode"])
    c = lambda a: a[..., 0] + 1j * a[..., 1]
    b32, b64 = c(res[hm.FP32]["bk"]), c(res[hm.FP64]["bk"])
    db = np.abs(b32 - b64).max() / np.abs(b64).max()
    report_err(f"metric B_{ells} fp32 vs fp64", db)
    assert db <= GEN_TOL


@pytest.mark.parametrize("ells", [(0, 0, 0), (1, 1, 0)])
def test_metric_zeta_fp32_vs_fp64(ctx, catalogue, ells):
    d, _ = catalogue
    rw = dict(n_rbin=10, rmin=20.0, rmax=200.0)
    res = {}
    for prec in (hm.FP32, hm.FP64):
        plan = hm.BoxPlan(ctx, N_MESH, BOX, "TSC", pk_ells=(), bk=False, zeta=True, zeta_ells=ells, precision=prec,
                          **rw)
        res[prec] = plan.execute(d)
        plan.close()
    c = lambda a: a[..., 0] + 1j * a[..., 1]
    z32, z64 = c(res[hm.FP32]["zeta"]), c(res[hm.FP64]["zeta"])
    assert np.isfinite(z64).all() and np.abs(z64).max() > 0
    dz = np.abs(z32 - z64).max() / np.abs(z64).max()
    report_err(f"metric zeta_{ells} fp32 vs fp64", dz)
    assert dz <= ZETA_TOL
