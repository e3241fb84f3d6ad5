"""Pins the oracle's learned-reconstruction network (no counterpart in the
reference):

  - the periodic convolution equals scipy.ndimage.correlate with mode="wrap"
    summed over input channels (a textbook routine);
  - the input and weight gradients are the adjoints of the convolution
    (<conv(x), y> = <x, input_grad(y)> = <W, weight_grad(x, y)>);
  - the FP64 network gradient (no rounding) equals central finite
    differences of sum(out * gout);
  - BF16 rounding is round-to-nearest-even on known bit patterns.
"""
import numpy as np
import pytest
from scipy import ndimage

from oracle import pyoracle as po


def rand_params(rng, scale=0.3):
    n = sum(27 * a * b + b for a, b in zip(po.LRECON_CIN, po.LRECON_COUT))
    return rng.normal(0.0, scale, n)


def test_bf16_round_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -2.5, 3.140625, 1e-3], dtype=np.float32)
    got = po.bf16_round(x)
    # 1 + 2^-8 is the tie between 1 and 1 + 2^-7: to even (1); 1 + 1.5 * 2^-8 rounds up
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0078125 and got[3] == 1.0078125
    assert got[4] == -2.5 and got[5] == 3.140625
    assert (got.view(np.uint32) & 0xFFFF).max() == 0
    assert abs(got[6] - 1e-3) <= 2 ** -9 * 1e-3


def test_conv_matches_scipy_correlate_wrap():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(6, 5, 7, 3))
    W = rng.normal(size=(27, 3, 2))
    y = po.conv3_periodic(x, W)
    for co in range(2):
        ref = np.zeros(x.shape[:3])
        for ci in range(3):
            K = W[:, ci, co].reshape(3, 3, 3)
            ref += ndimage.correlate(x[..., ci], K, mode="wrap")
        assert np.abs(y[..., co] - ref).max() < 1e-12


def test_conv_adjoints():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(5, 6, 4, 4))
    W = rng.normal(size=(27, 4, 3))
    g = rng.normal(size=(5, 6, 4, 3))
    lhs = np.sum(po.conv3_periodic(x, W) * g)
    assert abs(lhs - np.sum(x * po.conv3_input_grad(g, W))) < 1e-9 * abs(lhs)
    dW, db = po.conv3_weight_grad(x, g)
    assert abs(lhs - np.sum(W * dW)) < 1e-9 * abs(lhs)
    assert np.allclose(db, g.reshape(-1, 3).sum(0))


def test_network_gradient_finite_differences():
    rng = np.random.default_rng(3)
    n = 6
    x = rng.normal(size=(n, n, n, 4))
    p = rand_params(rng)
    gout = rng.normal(size=(n, n, n, 3))
    grad = po.lrecon_net_grad(x, p, gout, bf16=False)
    f = lambda q: np.sum(po.lrecon_net(x, q, bf16=False)[0] * gout)
    idx = rng.choice(p.size, 40, replace=False)
    # include the biases and the first and last layers
    idx = np.concatenate([idx, [0, 1740, p.size - 1, p.size - 5]])
    for i in idx:
        h = 1e-6
        e = np.zeros_like(p)
        e[i] = h
        fd = (f(p + e) - f(p - e)) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-5 * (1.0 + abs(fd)), i
