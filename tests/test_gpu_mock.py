"""GPU mock generator against the oracle: the Philox noise (counter-based
integers, FP64 Box-Muller) to 1e-13, and the Zel'dovich positions (GPU FP64
FFTs against numpy) to 1e-8 Mpc/h, for the whole lattice and for seeded
subsets of it, in real and redshift space."""
import numpy as np
import pytest

import hitomi_amd as hm
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

PK = dict(amp=2.0e4, k0=0.05, slope=-1.5, kcut=0.5)


@pytest.mark.parametrize("n,seed", [(16, 1), (32, 123456789012345)])
def test_white_noise_matches_oracle(ctx, n, seed):
    cfg = hm.mock_config(n, 500.0, **PK)
    got = hm.mock_white_noise(ctx, cfg, seed).download()
    ref = po.mock_white_noise(n, seed)
    assert np.abs(got - ref).max() <= 1e-13


@pytest.mark.parametrize("n,L,seed,frac,f,bao", [
    (32, 500.0, 5, 1.0, 0.0, 0.0),
    (32, 500.0, 6, 0.61, 0.0, 0.05),
    (64, 800.0, 7, 0.37, 0.76, 0.05),
])
def test_zeldovich_positions_match_oracle(ctx, n, L, seed, frac, f, bao):
    npart = int(frac * n ** 3)
    cfg = hm.mock_config(n, L, bao_amp=bao, f_growth=f, **PK)
    got = hm.mock_zeldovich(ctx, cfg, seed, npart).download()
    ref = po.mock_zeldovich(n, L, seed, npart, f_growth=f, bao_amp=bao, **PK)
    assert got.shape == ref.shape
    assert got.min() >= 0.0 and got.max() < L
    d = np.abs(got - ref)
    d = np.minimum(d, L - d)  # a coordinate within rounding of L may wrap on one side only
    assert d.max() <= 1e-8
    # the displacement is not trivially small
    lattice = (po.mock_sites(npart, n ** 3, seed)[:, None] // np.array([n * n, n, 1]) % n + 0.5) * (L / n)
    disp = np.abs(ref - lattice)
    disp = np.minimum(disp, L - disp)
    assert disp.std() > 0.05 * L / n


def test_rejects_bad_configuration(ctx):
    with pytest.raises(hm.HitomiError):
        hm.mock_zeldovich(ctx, hm.mock_config(48, 500.0, **PK), 1, 100)  # not a power of two
    with pytest.raises(hm.HitomiError):
        hm.mock_zeldovich(ctx, hm.mock_config(16, 500.0, **PK), 1, 16 ** 3 + 1)
