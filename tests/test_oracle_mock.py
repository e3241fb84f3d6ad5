"""Pins the oracle's mock generator (no counterpart in the reference):

  - Philox4x32-10 against the Random123 known-answer vectors;
  - the noise is counter-based (cell i's value depends on i and the seed
    only) and standard normal (moments and a Kolmogorov-Smirnov test);
  - the site choice is a bijection onto [0, sites), nested in n;
  - the Zel'dovich displacement of a single plane wave is the closed form
    Psi = -eps k / k^2 sin(k . x) (times 1 + f along z);
  - the Gaussian field's shell power recovers P(k) within cosmic variance.
"""
import numpy as np
import pytest
from scipy import stats

from oracle import pyoracle as po

# Random123 kat_vectors, philox4x32 with 10 rounds: (counter, key, output)
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", PHILOX_KAT)
def test_philox_known_answers(ctr, key, out):
    assert po.philox4x32_10(ctr, key) == out


def test_noise_counter_based_and_seeded():
    a = po.mock_white_noise(16, 11)
    b = po.mock_white_noise(32, 11)
    assert np.array_equal(a.ravel(), b.ravel()[: 16 ** 3])
    c = po.mock_white_noise(16, 12)
    assert not np.allclose(a, c)
    assert abs(np.corrcoef(a.ravel(), c.ravel())[0, 1]) < 0.1


def test_noise_standard_normal():
    z = po.mock_white_noise(64, 2024).ravel()
    n = z.size
    assert abs(z.mean()) < 5.0 / np.sqrt(n)
    assert abs(z.var() - 1.0) < 5.0 * np.sqrt(2.0 / n)
    assert abs(stats.skew(z)) < 5.0 * np.sqrt(6.0 / n)
    assert abs(stats.kurtosis(z)) < 5.0 * np.sqrt(24.0 / n)
    assert stats.kstest(z, "norm").pvalue > 1e-3
    # neighbouring cells are uncorrelated
    assert abs(np.corrcoef(z[:-1], z[1:])[0, 1]) < 5.0 / np.sqrt(n)


@pytest.mark.parametrize("sites", [1000, 4096, 27000, 32768])
def test_sites_bijection_and_nested(sites):
    s = po.mock_sites(sites - 1, sites, 77)
    assert s.min() >= 0 and s.max() < sites
    assert len(np.unique(s)) == sites - 1
    short = po.mock_sites(sites // 3, sites, 77)
    assert np.array_equal(short, s[: sites // 3])
    assert np.array_equal(po.mock_sites(sites, sites, 77), np.arange(sites))
    assert not np.array_equal(po.mock_sites(sites // 3, sites, 78), short)


@pytest.mark.parametrize("m,f", [((0, 0, 3), 0.0), ((2, -1, 1), 0.0), ((1, 2, 2), 0.76)])
def test_single_mode_zeldovich_closed_form(m, f):
    n, L, eps = 32, 400.0, 0.3
    x = (np.arange(n) * (L / n))
    X = np.stack(np.meshgrid(x, x, x, indexing="ij"), -1)
    k = 2.0 * np.pi * np.asarray(m, float) / L
    delta = eps * np.cos(X @ k)
    psi = po.zeldovich_displacement(np.fft.rfftn(delta), n, L, f_growth=f)
    s = np.sin(X @ k)
    k2 = k @ k
    for a in range(3):
        expect = -eps * k[a] / k2 * s * ((1.0 + f) if a == 2 else 1.0)
        assert np.abs(psi[a] - expect).max() < 1e-12 * eps / np.sqrt(k2) + 1e-15


def test_field_recovers_power_spectrum():
    n, L = 64, 1000.0
    kw = dict(amp=2.0e4, k0=0.05, slope=-1.5, kcut=0.5, bao_amp=0.05, bao_period=0.0628)
    dk = po.mock_delta_k(n, L, 99, kw)
    kx, ky, kz, nyq = po.mock_kgrid(n, L)
    kmag = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    p = np.abs(dk) ** 2 * L ** 3 / n ** 6
    kf = 2.0 * np.pi / L
    edges = kf * (np.arange(2, 26) + 0.5)
    sel = (~nyq) & (kz > 0)  # independent modes (kz = 0 plane holds conjugate pairs)
    chi = []
    for lo, hi in zip(edges[:-1], edges[1:]):
        s = sel & (kmag >= lo) & (kmag < hi)
        ratio = p[s] / po.mock_pk(kmag[s], **kw)
        chi.append((ratio.mean() - 1.0) * np.sqrt(s.sum()))
    chi = np.array(chi)
    assert np.abs(chi).max() < 5.0
    assert 0.3 < np.mean(chi ** 2) < 2.0


def test_sky_randoms_inside_patch_and_uniform():
    """The sky-patch randoms: every point inside the patch; r^3, RA and sin(Dec) uniform (Kolmogorov-Smirnov)
    and mutually uncorrelated; the count per volume element is the patch's mean density."""
    from scipy import stats

    from hitomi_amd import mocks
    import hitomi_amd as hm

    patch = hm.sky_patch(1200.0, 1800.0, 0.0, 60.0, -30.0, 30.0)
    x = po.sky_randoms([1200.0, 1800.0, 0.0, 60.0, -30.0, 30.0], 99, 200_000)
    assert mocks.in_sky_patch(x, patch).all()
    r = np.sqrt(np.sum(x * x, axis=1))
    u1 = (r ** 3 - 1200.0 ** 3) / (1800.0 ** 3 - 1200.0 ** 3)
    u2 = np.degrees(np.arctan2(x[:, 1], x[:, 0])) / 60.0
    s = x[:, 2] / r
    u3 = (s - np.sin(np.radians(-30.0))) / (np.sin(np.radians(30.0)) - np.sin(np.radians(-30.0)))
    for u in (u1, u2, u3):
        assert stats.kstest(u, "uniform").pvalue > 1e-3
    c = np.corrcoef(np.stack([u1, u2, u3]))
    assert np.abs(c[np.triu_indices(3, 1)]).max() < 0.01
    # an inner sub-shell holds its volume fraction of the points
    inner = (r < 1400.0).mean()
    assert inner == pytest.approx((1400.0 ** 3 - 1200.0 ** 3) / (1800.0 ** 3 - 1200.0 ** 3), abs=0.005)
    assert mocks.sky_patch_volume(patch) == pytest.approx(np.pi / 3 * (1800.0 ** 3 - 1200.0 ** 3) / 3, rel=1e-12)


def test_sky_randoms_seeded_and_counter_based():
    a = po.sky_randoms([500.0, 800.0, 0.0, 40.0, -20.0, 20.0], 5, 1000)
    b = po.sky_randoms([500.0, 800.0, 0.0, 40.0, -20.0, 20.0], 5, 300)
    c = po.sky_randoms([500.0, 800.0, 0.0, 40.0, -20.0, 20.0], 6, 300)
    assert np.array_equal(a[:300], b)
    assert not np.allclose(b, c)


def test_write_xyzw_format(tmp_path):
    """np.savetxt(..., fmt="%.10e") of x, y, z, w (fits2dat_galaxy_weight.py:134, 158)."""
    from hitomi_amd import io as hio

    pos = np.array([[1.0, -2.5e3, 3.25e-4], [1234.5678901234, 0.0, -7.0]])
    w = np.array([0.125, 1.0 / 3.0])
    f = tmp_path / "cat.dat"
    hio.write_xyzw(f, pos, w)
    lines = f.read_text().splitlines()
    assert lines[0] == "1.0000000000e+00 -2.5000000000e+03 3.2500000000e-04 1.2500000000e-01"
    assert lines[1] == "1.2345678901e+03 0.0000000000e+00 -7.0000000000e+00 3.3333333333e-01"
    back = np.loadtxt(f)
    assert np.allclose(back[:, :3], pos, rtol=1e-10) and np.allclose(back[:, 3], w, rtol=1e-10)
