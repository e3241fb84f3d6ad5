"""HIP Stockham FFT (hm_fft_c2c / r2c / c2r) against numpy.fft -- the
unnormalised FFTW convention of fftw_plan_dft_3d (field.hpp:321, 339)."""
import numpy as np
import pytest

import hitomi_amd as hm

pytestmark = pytest.mark.gpu

POW2 = [8, 16, 32, 64, 128, 256, 512, 1024]
MIXED = [24, 40, 48, 72, 80, 96, 120, 144, 160, 192, 200, 216, 240, 288, 320, 360, 384, 400, 432, 480, 576, 600, 640,
         720, 768, 800, 960]
# lengths with no compile-time plan: the run-time radix plan (radix 2-23 butterflies, a direct DFT stage for
# larger prime factors), including every axis length of the reference's BOSS meshes (RUN_boss.py:85-120)
RUNTIME = [2, 3, 5, 7, 11, 13, 17, 19, 23, 25, 26, 29, 46, 49, 97, 105, 130, 140, 169, 180, 190, 210, 250, 260, 280,
           340, 361, 460, 500, 529, 650, 1000, 1021, 1290, 2048]
TOL = {hm.FP64: 1e-12, hm.FP32: 2e-6}


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("prec", [hm.FP64, hm.FP32])
@pytest.mark.parametrize("L", POW2 + MIXED + [270] + RUNTIME)
def test_c2c_each_length_each_axis(ctx, prec, L):
    rng = np.random.default_rng(L)
    m = 8 if L <= 1024 else 4
    for shape in [(L, m, m), (m, L, m), (m, m, L)]:
        a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        f = ctx.fft_c2c(a, -1, prec)
        assert _rel(f, np.fft.fftn(a)) < TOL[prec] * np.log2(L), (shape, prec)
        b = ctx.fft_c2c(a, +1, prec)
        assert _rel(b, np.fft.ifftn(a) * a.size) < TOL[prec] * np.log2(L), (shape, prec)


@pytest.mark.parametrize("prec", [hm.FP64, hm.FP32])
@pytest.mark.parametrize("nz", [16, 32, 64, 128, 256, 512, 1024, 10, 26, 46, 130, 210, 260, 280, 360, 500, 540, 650,
                                2048])
def test_r2c_c2r_against_numpy(ctx, prec, nz):
    rng = np.random.default_rng(nz)
    a = rng.standard_normal((16, 8, nz))
    h = ctx.fft_r2c(a, prec)
    ref = np.fft.rfftn(a)
    assert h.shape == ref.shape
    assert _rel(h, ref) < TOL[prec] * np.log2(nz)
    back = ctx.fft_c2r(ref, nz, prec)
    assert _rel(back, np.fft.irfftn(ref, s=a.shape) * a.size) < TOL[prec] * np.log2(nz)


def test_c2c_mesh_512_round_trip(ctx):
    """Size-independent property at the metric mesh: backward(forward(x)) = N x."""
    rng = np.random.default_rng(512)
    a = (rng.standard_normal((512, 512, 512), dtype=np.float32)).astype(np.float32)
    h = ctx.fft_r2c(a, hm.FP32)
    back = ctx.fft_c2r(h, 512, hm.FP32) / a.size
    assert np.abs(back - a).max() < 5e-6 * np.abs(a).max()
    # Parseval on the half spectrum
    w = np.full(h.shape[2], 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    e_k = np.sum(w * np.sum(np.abs(h.astype(np.complex128)) ** 2, axis=(0, 1)))
    assert e_k == pytest.approx(a.size * np.sum(a.astype(np.float64) ** 2), rel=1e-5)
