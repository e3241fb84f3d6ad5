"""HIP Stockham FFT (hm_fft_c2c / r2c / c2r) against numpy.fft -- the
unnormalised FFTW convention of fftw_plan_dft_3d (field.hpp:321, 339)."""
import numpy as np
import pytest

import hitomi_amd as hm

pytestmark = pytest.mark.gpu

POW2 = [8, 16, 32, 64, 128, 256, 512, 1024]
MIXED = [24, 40, 48, 72, 80, 96, 120, 144, 160, 192, 200, 216, 240, 288, 320, 360, 384, 400, 432, 480, 576, 600, 640,
         720, 768, 800, 960]
TOL = {hm.FP64: 1e-12, hm.FP32: 2e-6}


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("prec", [hm.FP64, hm.FP32])
@pytest.mark.parametrize("L", POW2 + MIXED)
def test_c2c_each_length_each_axis(ctx, prec, L):
    rng = np.random.default_rng(L)
    for shape in [(L, 8, 8), (8, L, 8), (8, 8, L)]:
        a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        f = ctx.fft_c2c(a, -1, prec)
        assert _rel(f, np.fft.fftn(a)) < TOL[prec] * np.log2(L), (shape, prec)
        b = ctx.fft_c2c(a, +1, prec)
        assert _rel(b, np.fft.ifftn(a) * a.size) < TOL[prec] * np.log2(L), (shape, prec)


@pytest.mark.parametrize("prec", [hm.FP64, hm.FP32])
@pytest.mark.parametrize("nz", [16, 32, 64, 128, 256, 512, 1024])
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
