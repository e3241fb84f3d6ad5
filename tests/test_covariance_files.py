"""The covariance stage's file layout (hitomi_amd.covariance), on the CPU:
the readers take the columns calc_data_and_cov.py takes (xi: column 1 of
xi%d.dat and xi0_window.dat; zeta: column 2 of zeta%d%d%d_%02d.dat and
zeta000_%02d_window.dat, rows ith_rbin, flattened row-major; r bins from
column 0 / column 1, :66-123), from files in the drop-in programs' formats."""
import os

import numpy as np

from hitomi_amd import covariance as cv


def write_xi(d, ell, r, xi, suffix=""):
    with open(os.path.join(d, f"xi{ell}{suffix}.dat"), "w") as f:
        f.write("## r [Mpc/h] \t xi\n")
        for a, b in zip(r, xi):
            f.write("%.5f \t %.7e\n" % (a, b))


def write_zeta(d, l1, l2, L, r, z, suffix=""):
    for i in range(len(r)):
        with open(os.path.join(d, f"zeta{l1}{l2}{L}_{i:02d}{suffix}.dat"), "w") as f:
            f.write("## r1 [Mpc/h] \t r2 [Mpc/h] \t zeta \n")
            for j in range(len(r)):
                f.write("%.5f \t %.5f \t %.7e\n" % (r[i], r[j], z[i, j]))


def test_readers_take_the_reference_columns(tmp_path):
    rng = np.random.default_rng(3)
    r2 = np.linspace(30.0, 150.0, 25)
    r3 = np.linspace(20.0, 140.0, 13)
    d = str(tmp_path)
    xi = {l: rng.standard_normal(25) for l in cv.XI_ELLS}
    z = {m: rng.standard_normal((13, 13)) for m in cv.ZETA_MULTIPOLES}
    for l in cv.XI_ELLS:
        write_xi(d, l, r2, xi[l])
    for m in cv.ZETA_MULTIPOLES:
        write_zeta(d, *m, r3, z[m])
    write_xi(d, 0, r2, xi[0] + 2.0, suffix="_window")
    write_zeta(d, 0, 0, 0, r3, z[(0, 0, 0)] + 3.0, suffix="_window")
    for l in cv.XI_ELLS:
        assert np.array_equal(cv.read_xi(d, l), np.array(["%.7e" % v for v in xi[l]], dtype=float))
    for m in cv.ZETA_MULTIPOLES:
        got = cv.read_zeta(d, *m, 13)
        assert got.shape == (13, 13)
        assert np.array_equal(got, np.array([["%.7e" % v for v in row] for row in z[m]], dtype=float))
    assert np.allclose(cv.read_xi_window(d), xi[0] + 2.0, rtol=1e-7)
    assert np.allclose(cv.read_zeta(d, 0, 0, 0, 13, window=True), z[(0, 0, 0)] + 3.0, rtol=1e-6, atol=1e-6)
    rr2, rr3 = cv.read_rbins(d, 13)
    assert np.allclose(rr2, r2) and rr3.shape == (13, 13)
    # the flattening of calc_data_and_cov.py:98 and the r pairs of :125-126
    ri, rj = cv.zeta_rbins(rr3)
    assert np.allclose(ri, np.repeat(r3, 13)) and np.allclose(rj, np.tile(r3, 13))


def test_blocks_and_windows_layout():
    b = cv.blocks_for(25, 13)
    assert [n for n, _ in b] == ["xi0", "xi2", "xi4", "zeta000", "zeta110", "zeta202", "zeta112"]
    assert sum(s for _, s in b) == 3 * 25 + 4 * 169
    w0 = np.arange(1.0, 26.0)
    w000 = np.arange(1.0, 170.0).reshape(13, 13)
    w = cv.windows_for(w0, w000)
    assert all(np.array_equal(w[f"xi{l}"], w0) for l in (0, 2, 4))
    assert all(np.array_equal(w[cv.zeta_name(*m)], w000.ravel()) for m in cv.ZETA_MULTIPOLES)
