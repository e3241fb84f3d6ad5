"""Writes the golden vectors that pin the oracle's library restatements.

The reference ships no test vectors (its plot_test.py bodies are commented
out), so the values that pin the replaced FFTW/GSL routines are computed here
from independent implementations:
  - numpy.fft for the unnormalised 3D/1D DFT (fftw_plan_dft_3d semantics);
  - sympy exact rationals for the Wigner 3j symbols (gsl_sf_coupling_3j):
    every m for j1, j2 <= 4, and up to j1, j2 <= 8, j3 <= 16 (the multipoles
    the plans accept) every triangle with m = 0 plus a spread of m values;
  - sympy's exact associated Legendre functions, normalised as
    gsl_sf_legendre_sphPlm, for l <= 16;
  - mpmath at 30 digits for the spherical Bessel functions (gsl_sf_bessel_jl);
  - sympy's Ynm for the reduced spherical harmonics of tools.hpp:12-68, l <= 8.
Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os

import mpmath
import numpy as np
import sympy as sp
from sympy.physics.wigner import wigner_3j

HERE = os.path.dirname(os.path.abspath(__file__))


def fft_fixtures():
    rng = np.random.default_rng(20240611)
    out = {}
    for shape in [(8, 8, 8), (16, 16, 16), (10, 12, 14), (6, 9, 15), (4, 8, 46)]:
        a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        key = "x".join(map(str, shape))
        out[f"in_{key}"] = a
        out[f"fwd_{key}"] = np.fft.fftn(a)
        out[f"bwd_{key}"] = np.fft.ifftn(a) * a.size
    for n in [2, 3, 5, 7, 23, 60, 460]:
        a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"in1_{n}"] = a
        out[f"fwd1_{n}"] = np.fft.fft(a)
    np.savez_compressed(os.path.join(HERE, "fft_fixtures.npz"), **out)


def special():
    w3j = []
    seen = set()

    def add(j1, j2, j3, m1, m2, m3):
        key = (j1, j2, j3, m1, m2, m3)
        if key in seen or abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3 or m1 + m2 + m3 != 0:
            return
        seen.add(key)
        w3j.append([j1, j2, j3, m1, m2, m3, float(wigner_3j(j1, j2, j3, m1, m2, m3))])

    # every m for j1, j2 <= 4, j3 <= 6
    for j1 in range(0, 5):
        for j2 in range(0, 5):
            for j3 in range(abs(j1 - j2), min(j1 + j2, 6) + 1):
                for m1 in range(-j1, j1 + 1):
                    for m2 in range(-j2, j2 + 1):
                        add(j1, j2, j3, m1, m2, -m1 - m2)
    # the range the plans accept (ell1, ell2 <= 8, ELL <= 16): every triangle with m = 0, and for each a
    # spread of m1, m2 (the extremes, small values and one in between)
    for j1 in range(0, 9):
        for j2 in range(0, 9):
            for j3 in range(abs(j1 - j2), j1 + j2 + 1):
                for m1 in sorted({-j1, 0, 1, j1}):
                    for m2 in sorted({-j2, -1, 0, j2}):
                        add(j1, j2, j3, m1, m2, -m1 - m2)
    x = sp.Symbol("x")
    plm = []
    for l in range(0, 17):
        for m in range(0, l + 1):
            # P_l^m with the Condon-Shortley phase, then the sphPlm normalisation
            P = sp.assoc_legendre(l, m, x)
            norm = sp.sqrt(sp.Rational(2 * l + 1, 4) / sp.pi * sp.factorial(l - m) / sp.factorial(l + m))
            for xv in [-0.99, -0.7, -0.25, 0.0, 0.1, 0.5, 0.83, 1.0]:
                val = float(sp.N(norm * P.subs(x, sp.Rational(str(xv))), 30))
                plm.append([l, m, xv, val])
    mpmath.mp.dps = 30
    jl = []
    for l in range(0, 5):
        for xv in [0.0, 1e-6, 1e-3, 0.05, 0.5, 1.0, 2.0, 3.7, 5.0, 6.3, 10.0, 33.3, 100.0, 999.9, 9876.5]:
            if xv == 0.0:
                v = 1.0 if l == 0 else 0.0
            else:
                v = float(mpmath.sqrt(mpmath.pi / (2 * mpmath.mpf(xv))) * mpmath.besselj(l + 0.5, mpmath.mpf(xv)))
            jl.append([l, xv, v])
    # reduced Y: conj(Y_lm(theta, phi)) * sqrt(4 pi / (2l+1)) (tools.hpp:56-63)
    ylm = []
    th, ph = sp.symbols("theta phi", real=True)
    vecs = [(0.3, -0.4, 0.5), (1.0, 2.0, -3.0), (-0.2, 0.1, 0.9), (0.0, 0.0, 1.0), (0.6, -0.8, 0.0), (-1.0, -1.0, 1.0)]
    for l in range(0, 9):
        for m in range(-l, l + 1):
            Y = sp.Ynm(l, m, th, ph).expand(func=True)
            for v in vecs:
                r = float(np.sqrt(v[0] ** 2 + v[1] ** 2 + v[2] ** 2))
                tv = float(np.arccos(v[2] / r))
                pv = float(np.arctan2(v[1], v[0]))
                val = complex(sp.N(Y.subs({th: tv, ph: pv}), 25))
                val = val.conjugate() * float(np.sqrt(4 * np.pi / (2 * l + 1)))
                ylm.append([l, m, list(v), val.real, val.imag])
    with open(os.path.join(HERE, "special.json"), "w") as f:
        json.dump({"wigner3j": w3j, "sphPlm": plm, "bessel_jl": jl, "ylm_reduced": ylm}, f, separators=(",", ":"))


if __name__ == "__main__":
    fft_fixtures()
    special()
    print("wrote", os.listdir(HERE))
