"""Host-side special functions of libhitomi_amd.so (no GPU needed):
hm_sph_bessel_jl (the gsl_sf_bessel_jl node values of bessel.hpp:20)
against the mpmath golden values and scipy, and the Y_lm / 3j inputs that the
GPU estimators take from the library's own code are covered by the GPU
parity tests."""
import json
import os

import numpy as np
import pytest
from scipy.special import spherical_jn

import hitomi_amd as hm

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_sph_bessel_against_mpmath_golden():
    lib = hm.load_library()
    with open(os.path.join(GOLD, "special.json")) as f:
        special = json.load(f)
    for l, x, v in special["bessel_jl"]:
        assert lib.hm_sph_bessel_jl(int(l), float(x)) == pytest.approx(v, rel=1e-12, abs=1e-16), (l, x)


@pytest.mark.parametrize("l", range(0, 11))
def test_sph_bessel_against_scipy(l):
    lib = hm.load_library()
    xs = np.concatenate([np.linspace(0.0, 30.0, 3001), np.geomspace(30.0, 1e4, 500), [1e-9, 0.999999, 1.000001]])
    got = np.array([lib.hm_sph_bessel_jl(l, float(x)) for x in xs])
    ref = spherical_jn(l, xs)
    assert np.all(np.abs(got - ref) <= 1e-13 * np.maximum(np.abs(ref), 1e-2))
    assert np.isnan(lib.hm_sph_bessel_jl(l, -1.0)) and np.isnan(lib.hm_sph_bessel_jl(-1, 1.0))
