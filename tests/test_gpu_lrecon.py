"""Learned reconstruction on the GPU against the oracle:

  - the network's forward pass and parameter gradient (BF16 MFMA, FP32
    accumulation) against the oracle's BF16-emulating FP64 restatement, for
    meshes that use 1, 2 and 4 tiles per wave;
  - the features (plain reconstruction on the mesh) against the oracle's
    recon_psi and mass assignment;
  - the loss and its gradient on the output mesh (the fixed-point TSC
    scatter) against numpy;
  - the biased-tracer generator against the oracle;
  - training on generated mocks beats the plain reconstruction on held-out
    mocks (displacement error against the truth).
"""
import numpy as np
import pytest

import hitomi_amd as hm
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def he_params(rng, last_scale=0.05):
    parts = []
    for l, (ci, co) in enumerate(zip(po.LRECON_CIN, po.LRECON_COUT)):
        sd = np.sqrt(2.0 / (27 * ci)) if l < 3 else last_scale
        parts += [rng.normal(0.0, sd, 27 * ci * co), rng.normal(0.0, 0.05, co)]
    return np.concatenate(parts).astype(np.float32)


@pytest.mark.parametrize("n", [16, 32, 64])
def test_network_forward_and_gradient_match_oracle(ctx, n):
    rng = np.random.default_rng(n)
    m = hm.LRecon(ctx, n, 500.0)
    p = he_params(rng)
    m.set_params(p)
    assert np.array_equal(m.get_params(), p)
    x = rng.normal(0.0, 1.0, (n, n, n, 4)).astype(np.float32)
    out = m.net_forward(x)[..., :3]
    ref, _ = po.lrecon_net(x, p, bf16=True)
    scale = np.abs(ref).max()
    assert np.abs(out - ref).max() <= 1e-2 * scale
    assert np.sqrt(np.mean((out - ref) ** 2)) <= 1e-3 * scale
    gout = np.zeros((n, n, n, 4), np.float32)
    gout[..., :3] = rng.normal(0.0, 1.0, (n, n, n, 3))
    g = m.net_backward(x, gout)
    gref = po.lrecon_net_grad(x, p, gout[..., :3], bf16=True)
    o = 0
    for ci, co in zip(po.LRECON_CIN, po.LRECON_COUT):
        for size in (27 * ci * co, co):
            a, b = g[o:o + size], gref[o:o + size]
            assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-6, (ci, co, size)
            o += size
    m.close()


def test_features_match_oracle(ctx):
    from .conftest import clustered_catalogue

    n, L, b1, RG = 32, 600.0, 2.0, 15.0
    pos = clustered_catalogue(20000, L, seed=3, nclump=20, sigma=30.0)
    m = hm.LRecon(ctx, n, L, b1=b1, RG=RG, precision=hm.FP64)
    feat = m.features(ctx.array(pos)).download()
    p = po.param(n, L, po.TSC)
    field = po.calc_field(p, pos)
    psi = po.recon_psi(p, po.assignment_correction(p, po.fourier_transform(p, field)), b1, RG)
    norm = L ** 3 / len(pos)
    for a in range(3):
        ref = norm * psi[a].real
        assert np.abs(feat[..., a] - ref).max() <= 1e-6 * np.abs(ref).max()
    # counts per cell (the unwrapped stencil of the reference drops what falls past the last cell)
    count = (field * (L ** 3 / n ** 3)).real
    ref3 = count * n ** 3 / len(pos) - 1.0
    assert np.abs(feat[..., 3] - ref3).max() <= 1e-5 * np.abs(ref3).max()
    m.close()


def test_loss_and_output_gradient(ctx):
    """The plain shift is the reference's reconstruction gather of the features
    (particle_reconstruction.hpp:334-388, via the oracle), and the loss
    gradient on the output mesh is the reference's mass assignment (the
    oracle's calc_field) of the scaled residuals: the gather's adjoint."""
    n, L = 32, 500.0
    cfg = hm.mock_config(64, L, f_growth=0.76)
    dpos, dshift, N = hm.mock_zeldovich_biased(ctx, cfg, 11, 20000, 1.0)
    m = hm.LRecon(ctx, n, L)
    rng = np.random.default_rng(5)
    m.set_params(he_params(rng))
    l0, l1, grad, gout = m.gradient(dpos, dshift, want_gout=True)
    sp, sq = m.predict(dpos)
    pos, T = dpos.download(), dshift.download()
    S0, S1 = sp.download(), sq.download()
    assert abs(l0 - np.mean(np.sum((S0 - T) ** 2, 1))) <= 1e-12 * l0
    assert abs(l1 - np.mean(np.sum((S1 - T) ** 2, 1))) <= 1e-12 * l1
    p = po.param(n, L, po.TSC)
    feat = m.features(dpos).download()
    psi = np.ascontiguousarray(np.moveaxis(feat[..., :3].astype(np.float64), -1, 0)).astype(np.complex128)
    gathered = po.recon_gather(p, psi, 1.0, pos) - pos
    assert np.abs(gathered - S0).max() <= 1e-9 * np.abs(S0).max()
    # d loss / d out = (2 psi_scale / N) sum_i W r_i, r = S - T
    coef = 2.0 * 10.0 / N
    r = (S1 - T) * coef
    dV = (L / n) ** 3
    ref = np.zeros((n, n, n, 3))
    for a, wgt in enumerate((np.stack([r[:, 0], r[:, 1]], 1), np.stack([r[:, 2], 0 * r[:, 2]], 1))):
        f = po.calc_field(p, pos, wgt) * dV
        if a == 0:
            ref[..., 0], ref[..., 1] = f.real, f.imag
        else:
            ref[..., 2] = f.real
    g = gout.download()[..., :3]
    assert np.abs(g - ref).max() <= 1e-6 * np.abs(ref).max()
    m.close()


def test_biased_generator_matches_oracle(ctx):
    n, L, mean_n, b = 32, 400.0, 5000, 1.0
    PK = dict(amp=2.0e4, k0=0.05, slope=-1.5, kcut=0.5)
    cfg = hm.mock_config(n, L, f_growth=0.76, **PK)
    dpos, dshift, N = hm.mock_zeldovich_biased(ctx, cfg, 21, mean_n, b)
    pos_ref, shift_ref, keep = po.mock_zeldovich_biased(n, L, 21, mean_n, b, f_growth=0.76, **PK)
    assert N == len(keep)
    assert abs(N - mean_n) < 5 * np.sqrt(mean_n)
    d = np.abs(dpos.download() - pos_ref)
    d = np.minimum(d, L - d)
    assert d.max() <= 1e-8
    assert np.abs(dshift.download() - shift_ref).max() <= 1e-8
    # the selection follows the linear field: kept sites are denser where delta_L > 0
    dk = po.mock_delta_k(n, L, 21, PK)
    _, _, _, nyq = po.mock_kgrid(n, L)
    dk = np.where(nyq, 0.0, dk)
    dk[0, 0, 0] = 0.0
    dl = np.fft.irfftn(dk, s=(n,) * 3, axes=(0, 1, 2)).ravel()
    assert dl[keep].mean() > 0.5 * dl.std() ** 2 / dl.std()


def test_training_beats_plain_reconstruction(ctx):
    n, L, side = 32, 500.0, 64
    cfg = hm.mock_config(side, L, f_growth=0.76, bao_amp=0.05)
    m = hm.LRecon(ctx, n, L, b1=2.0, RG=15.0, lr=3e-3, init_seed=7)
    train_seeds, test_seeds = range(100, 124), range(900, 906)
    mocks = {s: hm.mock_zeldovich_biased(ctx, cfg, s, 25000, 1.0) for s in list(train_seeds) + list(test_seeds)}
    before = [m.gradient(mocks[s][0], mocks[s][1])[:2] for s in test_seeds]
    assert all(abs(a - b) <= 1e-12 * a for a, b in before)  # a new model predicts the plain shifts
    for epoch in range(3):
        for s in train_seeds:
            m.train_step(mocks[s][0], mocks[s][1])
    after = np.array([m.gradient(mocks[s][0], mocks[s][1])[:2] for s in test_seeds])
    plain, learned = after[:, 0], after[:, 1]
    print("held-out displacement rms: plain", np.sqrt(plain.mean()), "learned", np.sqrt(learned.mean()))
    assert np.all(learned < plain)
    assert learned.mean() < 0.9 * plain.mean()
    m.close()


@pytest.mark.parametrize("n", [32, 64])
def test_lds_staged_convolutions_equal_cache_reads(ctx, n):
    """The LDS-staged convolution kernels give the cache-reading ones' results bit for bit."""
    rng = np.random.default_rng(100 + n)
    p = he_params(rng)
    x = rng.normal(0.0, 1.0, (n, n, n, 4)).astype(np.float32)
    gout = np.zeros((n, n, n, 4), np.float32)
    gout[..., :3] = rng.normal(0.0, 1.0, (n, n, n, 3))
    res = []
    for cache in (False, True):
        m = hm.LRecon(ctx, n, 500.0, conv_cache_taps=cache)
        m.set_params(p)
        res.append((m.net_forward(x), m.net_backward(x, gout)))
        m.close()
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])
