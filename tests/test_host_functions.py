"""Host-side special functions of libhitomi_amd.so (no GPU needed):
hm_sph_bessel_jl (the gsl_sf_bessel_jl node values of bessel.hpp:20)
against the mpmath golden values and scipy; hm_wigner_3j (gsl_sf_coupling_3j)
and hm_ylm_reduced (tools.hpp:12-68) against the sympy golden values."""
import ctypes as ct
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


def test_wigner3j_and_ylm_against_sympy_golden():
    """The library's own 3j (Racah formula in long double) and reduced Y_lm
    (the function the kernels evaluate) against the sympy fixtures, which
    cover j1, j2 <= 8, j3 <= 16 and l <= 8."""
    lib = hm.load_library()
    with open(os.path.join(GOLD, "special.json")) as f:
        special = json.load(f)
    for j1, j2, j3, m1, m2, m3, v in special["wigner3j"]:
        assert lib.hm_wigner_3j(j1, j2, j3, m1, m2, m3) == pytest.approx(v, abs=1e-14), (j1, j2, j3, m1, m2, m3)
    assert lib.hm_wigner_3j(8, 8, 17, 0, 0, 0) == 0.0  # outside the triangle
    assert abs(lib.hm_wigner_3j(3, 3, 3, 0, 0, 0)) <= 1e-14  # odd sum with zero m (zero up to rounding)
    out = (ct.c_double * 2)()
    for l, m, v, re, im in special["ylm_reduced"]:
        lib.hm_ylm_reduced(l, m, (ct.c_double * 3)(*v), out)
        assert out[0] == pytest.approx(re, abs=1e-12) and out[1] == pytest.approx(im, abs=1e-12), (l, m, v)
    lib.hm_ylm_reduced(0, 0, (ct.c_double * 3)(0.0, 0.0, 0.0), out)
    assert (out[0], out[1]) == (1.0, 0.0)
    lib.hm_ylm_reduced(2, 1, (ct.c_double * 3)(0.0, 0.0, 0.0), out)
    assert (out[0], out[1]) == (0.0, 0.0)
