"""The drop-in program hitomi_box (main_box.cpp): parameter-file semantics and
rejections on the CPU; end-to-end outputs against the oracle on the GPU."""
import os
import subprocess

import numpy as np
import pytest

from hitomi_amd import io as hio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "hitomi_amd", "bin", "hitomi_box")


def run(ini, cwd):
    if not os.path.exists(EXE):
        pytest.fail("hitomi_box is not built: run __graft_entry__.build()")
    return subprocess.run([EXE, str(ini)], cwd=cwd, capture_output=True, text=True, timeout=300)


def test_no_argument_exits_1(tmp_path):
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 1 and "The input parameter file is not found." in r.stdout


@pytest.mark.parametrize("key,value,message", [
    ("measure", "xi", "The 'measure' is not loaded correctly."),
    ("boxsize_y", "0", "The 'boxsize' is not loaded correctly."),
    ("n_mesh_z", "-4", "The 'n_mesh' is not loaded correctly."),
    ("ell2", "-1", "The 'ell1', 'ell2', or 'ELL' is not loaded correctly."),
    ("kmin", "0", "The 'kmin' is not loaded correctly."),
    ("kmax", "0.001", "The 'kmax' is not loaded correctly."),
    ("n_kbin", "1", "The 'n_kbin' is not loaded correctly."),
    ("ith_kbin", "20", "The 'ith_kbin' is not loaded correctly."),
    ("assignment", "PCS", "The 'assign' is not loaded correctly."),
    ("flag_recon", "Maybe", "The 'flag_recon' is not loaded correctly."),
    ("sim_data", "Tipsy", "The 'sim_data' is not loaded correctly."),
    ("log10_Mmax", "12.0", "The 'log10_Mmax' is not loaded correctly."),
    ("data_mpi_file", None, "The 'data_mpi_file' is not loaded correctly."),
])
def test_rejected_parameters(tmp_path, key, value, message):
    ini = tmp_path / "p.ini"
    hio.write_ini(ini, output_dir=str(tmp_path / "out"), **{key: value})
    r = run(ini, tmp_path)
    assert r.returncode == 1
    assert message in r.stdout and "'param.readParameters' fails." in r.stdout


def test_substring_keys(tmp_path):
    """parameter.hpp:161-207: keys match anywhere in a line.  A line whose
    *comment* names another key also sets that key (here RG is set from a line
    about NR), exactly as the reference parses it."""
    ini = tmp_path / "p.ini"
    hio.write_ini(ini, output_dir=str(tmp_path / "out"), RG=None, data_dir=str(tmp_path))
    with open(ini, "a") as f:
        f.write("NR = 0 %% also parsed as RG\n")
    (tmp_path / "snap.0").write_bytes(b"")
    r = run(ini, tmp_path)
    # the catalogue is empty, so the run stops at the reader -- after the parameters were accepted
    assert "'param.readParameters' fails." not in r.stdout
    txt = (tmp_path / "out" / "input_parameters").read_text()
    assert "RG = 0.00000\n" in txt and "NR = 0\n" in txt


def test_input_parameters_file(tmp_path):
    ini = tmp_path / "p.ini"
    out = tmp_path / "out"
    hio.write_ini(ini, output_dir=str(out), data_dir=str(tmp_path), realization=3, n_mesh_x=32, kmin=0.005)
    run(ini, tmp_path)
    txt = (out / "0003" / "input_parameters").read_text()
    assert txt.startswith(f"\ndata_dir = {tmp_path}\ndata_file = {tmp_path}/snap\nrandom_file = {tmp_path}/AAA\n"
                          f"output_dir = {out}/0003\n\nmeasure = pk\n\nrealization = 3\n")
    assert "n_mesh_x = 32\nn_mesh_y = 64\n" in txt and "kmin = 0.00500\n" in txt
    assert txt.endswith("NR = 0\ndata_mpi_file = mock\n\nassignment = TSC\n\n")


def test_missing_snapshot_exits_1(tmp_path):
    ini = tmp_path / "p.ini"
    hio.write_ini(ini, output_dir=str(tmp_path / "out"), data_dir=str(tmp_path), data_file="nothere")
    r = run(ini, tmp_path)
    assert r.returncode == 1 and "can not open" in r.stdout


@pytest.mark.gpu
def test_end_to_end_pk_and_bk_match_oracle(tmp_path):
    from oracle import pyoracle as po
    from .conftest import clustered_catalogue

    L, n, N = 1000.0, 32, 30000
    pos = clustered_catalogue(N, L, seed=77, nclump=20, sigma=25.0).astype(np.float32)
    hio.write_gadget1(str(tmp_path / "snap"), pos, boxsize=L, num_files=2)
    posd = pos.astype(np.float64)
    kw = dict(data_dir=str(tmp_path), boxsize_x=L, boxsize_y=L, boxsize_z=L, n_mesh_x=n, n_mesh_y=n, n_mesh_z=n,
              kmin=0.02, kmax=0.08, n_kbin=4, precision="FP64")
    for ell in (0, 2):
        ini = tmp_path / f"pk{ell}.ini"
        hio.write_ini(ini, output_dir=str(tmp_path / "out"), measure="pk", ell1=ell, ELL=ell, **kw)
        r = run(ini, tmp_path)
        assert r.returncode == 0, r.stdout
        got = hio.read_table(tmp_path / "out" / f"pk{ell}.dat")
        p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.08, n_kbin=4, ell1=ell, ELL=ell)
        ref, _ = po.box_pk(p, posd)
        assert np.allclose(got[:, 0], po.kbin(p), atol=5e-6)
        assert np.allclose(got[:, 1], ref.real, rtol=2e-7, atol=1e-7 * np.abs(ref.real).max())
    ini = tmp_path / "bk.ini"
    hio.write_ini(ini, output_dir=str(tmp_path / "out"), measure="bk", ith_kbin=2, **kw)
    r = run(ini, tmp_path)
    assert r.returncode == 0, r.stdout
    got = hio.read_table(tmp_path / "out" / "bk000_02.dat")
    p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.08, n_kbin=4, ith_kbin=2)
    ref, _, _ = po.box_bk(p, posd)
    assert np.allclose(got[:, 2], ref.real, rtol=1e-6, atol=1e-7 * np.abs(ref.real).max())
    assert np.allclose(got[:, 3], ref.imag, atol=1e-6 * np.abs(ref.real).max())
    assert not (tmp_path / "out" / "bk000_01.dat").exists()
    ini = tmp_path / "bk110.ini"
    hio.write_ini(ini, output_dir=str(tmp_path / "out"), measure="bk", ith_kbin=1, ell1=1, ell2=1, ELL=0, **kw)
    r = run(ini, tmp_path)
    assert r.returncode == 0, r.stdout
    got = hio.read_table(tmp_path / "out" / "bk110_01.dat")
    p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.08, n_kbin=4, ith_kbin=1, ell1=1, ell2=1, ELL=0)
    ref, _, _ = po.box_bk(p, posd)
    assert np.allclose(got[:, 2], ref.real, rtol=1e-6, atol=1e-7 * np.abs(ref.real).max())
    for ell in (0, 2):
        ini = tmp_path / f"xi{ell}.ini"
        hio.write_ini(ini, output_dir=str(tmp_path / "out"), measure="2PCF", ell1=ell, ELL=ell, rmin=40.0,
                      rmax=200.0, n_rbin=9, **kw)
        r = run(ini, tmp_path)
        assert r.returncode == 0, r.stdout
        got = hio.read_table(tmp_path / "out" / f"xi{ell}.dat")
        p = po.param(n, L, po.TSC, rmin=40.0, rmax=200.0, n_rbin=9, ell1=ell, ELL=ell)
        ref = po.box_xi(p, posd)
        assert np.allclose(got[:, 0], po.rbin(p), atol=5e-6)
        assert np.allclose(got[:, 1], ref, rtol=2e-7, atol=1e-7 * np.abs(ref).max())
    ini = tmp_path / "zeta.ini"
    hio.write_ini(ini, output_dir=str(tmp_path / "out"), measure="3PCF", ith_rbin=3, ell1=1, ell2=1, ELL=0,
                  rmin=40.0, rmax=200.0, n_rbin=6, **kw)
    r = run(ini, tmp_path)
    assert r.returncode == 0, r.stdout
    got = hio.read_table(tmp_path / "out" / "zeta110_03.dat")
    p = po.param(n, L, po.TSC, rmin=40.0, rmax=200.0, n_rbin=6, ith_rbin=3, ell1=1, ell2=1, ELL=0)
    ref, _, _ = po.box_zeta(p, posd)
    rb = po.rbin(p)
    assert np.allclose(got[:, 0], rb[3], atol=5e-6) and np.allclose(got[:, 1], rb, atol=5e-6)
    assert np.allclose(got[:, 2], ref.real, rtol=1e-6, atol=1e-7 * np.abs(ref.real).max())


SURVEY = os.path.join(ROOT, "hitomi_amd", "bin", "hitomi_survey")


def run_survey(ini, cwd):
    if not os.path.exists(SURVEY):
        pytest.fail("hitomi_survey is not built: run __graft_entry__.build()")
    return subprocess.run([SURVEY, str(ini)], cwd=cwd, capture_output=True, text=True, timeout=300)


def write_xyzw(path, pos, w):
    with open(path, "w") as f:
        f.write("# x y z w\n")
        for (x, y, z), ww in zip(pos, w):
            f.write(f"{x:.10e} {y:.10e} {z:.10e} {ww:.10e}\n")


def test_survey_missing_randoms_exits_1(tmp_path):
    write_xyzw(tmp_path / "data.dat", np.ones((3, 3)) * 500.0, np.ones(3))
    ini = tmp_path / "p.ini"
    hio.write_ini(ini, data_dir=str(tmp_path), data_file="data.dat", random_file="nothere.dat",
                  output_dir=str(tmp_path / "out"))
    r = run_survey(ini, tmp_path)
    assert r.returncode == 1 and "Cannot open the file" in r.stdout and "'P_R.readParticleBOSS' fails." in r.stdout


@pytest.mark.parametrize("measure", ["bk", "3PCF", "window3PCF"])
@pytest.mark.parametrize("ells", [(1, 0, 0), (2, 0, 1), (1, 1, 3)])
def test_survey_disallowed_multipole_exits_1(tmp_path, measure, ells):
    """bispec.hpp:12-18, 363-369, 624-630: wigner_3j(ell1, ell2, ELL, 0, 0, 0) = 0 is rejected with the
    reference's message before any GPU work."""
    write_xyzw(tmp_path / "data.dat", np.ones((3, 3)) * 500.0, np.ones(3))
    ini = tmp_path / "p.ini"
    hio.write_ini(ini, data_dir=str(tmp_path), data_file="data.dat", random_file="data.dat",
                  output_dir=str(tmp_path / "out"), measure=measure, ell1=ells[0], ell2=ells[1], ELL=ells[2])
    r = run_survey(ini, tmp_path)
    assert r.returncode == 1
    assert "This multipole combination is not allowed." in r.stdout
    assert "It should be wigner_3j(ell1, ell2, ELL, 0,0,0) != 0" in r.stdout


@pytest.mark.gpu
def test_survey_end_to_end_matches_oracle(tmp_path):
    from oracle import pyoracle as po

    rng = np.random.default_rng(5)
    lo, size = np.array([600.0, 900.0, 300.0]), 500.0
    posD = lo + rng.random((2500, 3)) * size
    posR = lo + rng.random((10000, 3)) * size
    wD = 1.0 / (1.0 + rng.random(2500))
    wR = 1.0 / (1.0 + rng.random(10000))
    write_xyzw(tmp_path / "data.dat", posD, wD)
    write_xyzw(tmp_path / "rand.dat", posR, wR)
    # what the files hold, as the reader parses them
    rd = lambda f: np.loadtxt(f, comments="#")
    D, R = rd(tmp_path / "data.dat"), rd(tmp_path / "rand.dat")
    n, L = 32, 700.0
    for ell in (0, 2):
        ini = tmp_path / f"pk{ell}.ini"
        hio.write_ini(ini, data_dir=str(tmp_path), data_file="data.dat", random_file="rand.dat",
                      output_dir=str(tmp_path / "out"), measure="pk", ell1=ell, ELL=ell, boxsize_x=L, boxsize_y=L,
                      boxsize_z=L, n_mesh_x=n, n_mesh_y=n, n_mesh_z=n, kmin=0.02, kmax=0.12, n_kbin=6,
                      precision="FP64")
        r = run_survey(ini, tmp_path)
        assert r.returncode == 0, r.stdout
        got = hio.read_table(tmp_path / "out" / f"pk{ell}.dat")
        p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.12, n_kbin=6, ell1=ell, ELL=ell)
        s = po.Survey(p, D[:, :3], D[:, 3], R[:, :3], R[:, 3])
        ref, _ = s.pk(p)
        assert np.allclose(got[:, 0], po.kbin(p), atol=5e-6)
        assert np.allclose(got[:, 1], ref.real, rtol=2e-7, atol=1e-7 * np.abs(ref.real).max())
        assert np.allclose(got[:, 2], ref.imag, rtol=2e-7, atol=1e-6 * np.abs(ref.real).max())
    for measure, fname in (("2PCF", "xi2.dat"), ("window2PCF", "xi2_window.dat")):
        ini = tmp_path / f"{measure}.ini"
        hio.write_ini(ini, data_dir=str(tmp_path), data_file="data.dat", random_file="rand.dat",
                      output_dir=str(tmp_path / "out"), measure=measure, ell1=2, ELL=2, boxsize_x=L, boxsize_y=L,
                      boxsize_z=L, n_mesh_x=n, n_mesh_y=n, n_mesh_z=n, rmin=40.0, rmax=200.0, n_rbin=5,
                      precision="FP64")
        r = run_survey(ini, tmp_path)
        assert r.returncode == 0, r.stdout
        got = hio.read_table(tmp_path / "out" / fname)
        p = po.param(n, L, po.TSC, rmin=40.0, rmax=200.0, n_rbin=5, ell1=2, ELL=2)
        s = po.Survey(p, D[:, :3], D[:, 3], R[:, :3], R[:, 3])
        ref, _ = s.xi(p, window=measure == "window2PCF")
        assert np.allclose(got[:, 0], po.rbin(p), atol=5e-6)
        assert np.allclose(got[:, 1], ref.real, rtol=2e-7, atol=1e-7 * np.abs(ref.real).max())
    txt = (tmp_path / "out" / "survey_parameters").read_text()
    assert f"P_D.n_tot = {len(D)}" in txt and f"P_R.n_tot = {len(R)}" in txt
    assert f"alpha = {s.alpha:.5f}" in txt
    assert f"nbar / 1.0e-4 [(Mpc/h)^{{-3}}] = {s.nbar / 1e-4:.5f}" in txt


@pytest.mark.gpu
def test_box_reconstruction_end_to_end(tmp_path):
    """main_box.cpp with flag_recon = True: randoms read, reconstruction on n_mesh_recon,
    wrap, P_l / xi_l of D - alpha R written to pk%d_recon.dat / xi%d_recon.dat."""
    from oracle import pyoracle as po
    from .conftest import clustered_catalogue, uniform_catalogue

    L, n, nrm, N, NR = 1000.0, 32, 24, 20000, 50000
    pos = clustered_catalogue(N, L, seed=81, nclump=20, sigma=25.0).astype(np.float32)
    hio.write_gadget1(str(tmp_path / "snap"), pos, boxsize=L, num_files=1)
    R = uniform_catalogue(NR, L, seed=82)
    with open(tmp_path / "rand.dat", "w") as f:
        for x, y, z in R:
            f.write(f"{x:.10e} {y:.10e} {z:.10e}\n")
    Rf = np.loadtxt(tmp_path / "rand.dat")
    b1, RG = 2.0, 20.0
    kw = dict(data_dir=str(tmp_path), random_file="rand.dat", boxsize_x=L, boxsize_y=L, boxsize_z=L, n_mesh_x=n,
              n_mesh_y=n, n_mesh_z=n, n_mesh_recon_x=nrm, n_mesh_recon_y=nrm, n_mesh_recon_z=nrm, flag_recon="True",
              b1_fid=b1, RG=RG, kmin=0.01, kmax=0.15, n_kbin=8, rmin=30.0, rmax=200.0, n_rbin=10, precision="FP64")
    dD, dR = po.recon_displace(po.param(nrm, L, po.TSC), pos.astype(np.float64), None, Rf, None, float(N // NR),
                               L ** 3 / N, b1, RG)
    for a in (dD, dR):
        for ax in range(3):
            a[a[:, ax] >= L, ax] -= L
            a[a[:, ax] < 0.0, ax] += L
    alpha = N / NR
    for measure, fname, xi in (("pk", "pk2_recon.dat", False), ("2PCF", "xi2_recon.dat", True)):
        ini = tmp_path / f"{measure}.ini"
        hio.write_ini(ini, output_dir=str(tmp_path / "out"), measure=measure, ell1=2, ELL=2, **kw)
        r = run(ini, tmp_path)
        assert r.returncode == 0, r.stdout
        got = hio.read_table(tmp_path / "out" / fname)
        p = po.param(n, L, po.TSC, ell1=2, ELL=2, kmin=0.01, kmax=0.15, n_kbin=8, rmin=30.0, rmax=200.0, n_rbin=10)
        ref, _ = po.box_recon_2pt(p, dD, dR, alpha, xi=xi)
        assert np.allclose(got[:, 1], ref.real, rtol=1e-6, atol=1e-7 * np.abs(ref.real).max())
    txt = (tmp_path / "out" / "simulation_parameters").read_text()
    assert f"P_R.n_tot = {NR}" in txt and f"alpha = {alpha:.5f}" in txt


@pytest.mark.gpu
def test_survey_reconstruction_end_to_end(tmp_path):
    from oracle import pyoracle as po

    rng = np.random.default_rng(9)
    lo, size = np.array([600.0, 900.0, 300.0]), 500.0
    posD = lo + rng.random((3000, 3)) * size
    posR = lo + rng.random((12000, 3)) * size
    wD = 1.0 / (1.0 + rng.random(3000))
    wR = 1.0 / (1.0 + rng.random(12000))
    write_xyzw(tmp_path / "data.dat", posD, wD)
    write_xyzw(tmp_path / "rand.dat", posR, wR)
    D, R = np.loadtxt(tmp_path / "data.dat", comments="#"), np.loadtxt(tmp_path / "rand.dat", comments="#")
    n, L, nrm, b1, RG = 32, 700.0, 24, 2.0, 20.0
    ini = tmp_path / "p.ini"
    hio.write_ini(ini, data_dir=str(tmp_path), data_file="data.dat", random_file="rand.dat",
                  output_dir=str(tmp_path / "out"), measure="pk", ell1=0, ELL=0, boxsize_x=L, boxsize_y=L,
                  boxsize_z=L, n_mesh_x=n, n_mesh_y=n, n_mesh_z=n, kmin=0.02, kmax=0.12, n_kbin=6, precision="FP64",
                  flag_recon="True", n_mesh_recon_x=nrm, n_mesh_recon_y=nrm, n_mesh_recon_z=nrm, b1_fid=b1, RG=RG)
    r = run_survey(ini, tmp_path)
    assert r.returncode == 0, r.stdout
    got = hio.read_table(tmp_path / "out" / "pk0_recon.dat")
    p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.12, n_kbin=6)
    s = po.Survey(p, D[:, :3], D[:, 3], R[:, :3], R[:, 3])
    dD, dR = po.recon_displace(po.param(nrm, L, po.TSC), s.posD, D[:, 3], s.posR, R[:, 3], s.alpha,
                               s.vsurvey / D[:, 3].sum(), b1, RG)
    d2 = dR.min(0) - L / n
    s.posD, s.posR = dD - d2, dR - d2
    ref, _ = s.pk(p)
    assert np.allclose(got[:, 1], ref.real, rtol=1e-6, atol=1e-7 * np.abs(ref.real).max())


@pytest.mark.gpu
def test_box_default_fp32_end_to_end(tmp_path):
    """A parameter file without a precision key runs FP32 grids and FFTs (the drop-in's default): pk and bk
    outputs against the oracle within the FP32 tolerances (P: 1e-6 of |P_0| + V/N; B: 5e-6 of max|B|).
    (Keys match anywhere in a line, as parameter.hpp:161-207 reads them, so a data_dir whose path holds the
    word of a key sets that key too: the test's name, part of its directory, avoids them.)"""
    from oracle import pyoracle as po
    from .conftest import clustered_catalogue

    L, n, N = 1000.0, 32, 30000
    pos = clustered_catalogue(N, L, seed=78, nclump=20, sigma=25.0).astype(np.float32)
    hio.write_gadget1(str(tmp_path / "snap"), pos, boxsize=L, num_files=1)
    posd = pos.astype(np.float64)
    kw = dict(data_dir=str(tmp_path), boxsize_x=L, boxsize_y=L, boxsize_z=L, n_mesh_x=n, n_mesh_y=n, n_mesh_z=n,
              kmin=0.02, kmax=0.08, n_kbin=4)
    ref0 = None
    for ell in (0, 2):
        ini = tmp_path / f"pk{ell}.ini"
        hio.write_ini(ini, output_dir=str(tmp_path / "out"), measure="pk", ell1=ell, ELL=ell, **kw)
        r = run(ini, tmp_path)
        assert r.returncode == 0, r.stdout
        got = hio.read_table(tmp_path / "out" / f"pk{ell}.dat")
        p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.08, n_kbin=4, ell1=ell, ELL=ell)
        ref, _ = po.box_pk(p, posd)
        ref0 = ref if ell == 0 else ref0
        # the file holds %.7e: 5e-8 relative of rounding on top of the FP32 error
        assert np.all(np.abs(got[:, 1] - ref.real) <= 1e-6 * (np.abs(ref0.real) + L ** 3 / N))
    ini = tmp_path / "bk.ini"
    hio.write_ini(ini, output_dir=str(tmp_path / "out"), measure="bk", ith_kbin=1, **kw)
    r = run(ini, tmp_path)
    assert r.returncode == 0, r.stdout
    got = hio.read_table(tmp_path / "out" / "bk000_01.dat")
    p = po.param(n, L, po.TSC, kmin=0.02, kmax=0.08, n_kbin=4, ith_kbin=1)
    ref, _, _ = po.box_bk(p, posd)
    assert np.abs(got[:, 2] - ref.real).max() <= 5e-6 * np.abs(ref.real).max()


@pytest.mark.gpu
def test_survey_boss_mesh_default_fp32_matches_oracle(tmp_path):
    """hitomi_survey on the reference's North zbin-1 BOSS box (1350 x 2450 x 1400 Mpc/h, RUN_boss.py:85-92)
    with its mesh divided by 10 (25 x 46 x 26: odd, non-power-of-two, anisotropic cells), a survey-geometry
    mock written as x y z w text, no precision key (FP32): pk0.dat and pk2.dat against the oracle within
    2e-5 of |P_0| + norm SN_00."""
    import hitomi_amd as hm
    from hitomi_amd import mocks
    from oracle import pyoracle as po

    n, L = (25, 46, 26), (1350.0, 2450.0, 1400.0)
    with hm.Context(0) as ctx:
        cat = mocks.survey_catalogue(ctx, hm.sky_patch(500.0, 800.0, 0.0, 40.0, -20.0, 20.0), 20_000, 100_000,
                                     seed=31, n_side=64, randoms_on_device=False)
    hio.write_xyzw(tmp_path / "data.dat", cat["posD"], cat["wD"])
    hio.write_xyzw(tmp_path / "rand.dat", cat["posR"], cat["wR"])
    D, R = np.loadtxt(tmp_path / "data.dat"), np.loadtxt(tmp_path / "rand.dat")
    kw = dict(kmin=0.01, kmax=0.1, n_kbin=10)
    p0 = po.param(n, L, po.TSC, **kw)
    s0 = po.Survey(p0, D[:, :3], D[:, 3], R[:, :3], R[:, 3])
    ref0, _ = s0.pk(p0)
    sn0 = s0.shotnoise(0, 0).real * s0.vsurvey / s0.wsumD ** 2
    for ell in (0, 2):
        ini = tmp_path / f"pk{ell}.ini"
        hio.write_ini(ini, data_dir=str(tmp_path), data_file="data.dat", random_file="rand.dat",
                      output_dir=str(tmp_path / "out"), measure="pk", ell1=ell, ELL=ell, boxsize_x=L[0],
                      boxsize_y=L[1], boxsize_z=L[2], n_mesh_x=n[0], n_mesh_y=n[1], n_mesh_z=n[2], **kw)
        r = run_survey(ini, tmp_path)
        assert r.returncode == 0, r.stdout
        got = hio.read_table(tmp_path / "out" / f"pk{ell}.dat")
        p = po.param(n, L, po.TSC, ell1=ell, ELL=ell, **kw)
        ref, _ = po.Survey(p, D[:, :3], D[:, 3], R[:, :3], R[:, 3]).pk(p)
        assert np.allclose(got[:, 0], po.kbin(p), atol=5e-6)
        assert np.all(np.abs(got[:, 1] - ref.real) <= 2e-5 * (np.abs(ref0.real) + sn0))


@pytest.mark.gpu
@pytest.mark.parametrize("mesh", [((250, 460, 260), (1350.0, 2450.0, 1400.0)),
                                  ((340, 650, 360), (1800.0, 3400.0, 1900.0)),
                                  ((190, 360, 210), (1000.0, 1900.0, 1100.0)),
                                  ((190, 500, 280), (1000.0, 2600.0, 1500.0))],
                         ids=["NGC_z1", "NGC_z3", "SGC_z1", "SGC_z3"])
def test_survey_accepts_boss_mesh_box_pairs(tmp_path, mesh):
    """Every mesh / box pair of RUN_boss.py:85-120 runs through hitomi_survey (pk, default FP32), and its
    pk0.dat agrees with an FP64 run within 1e-5 of |P_0|."""
    import hitomi_amd as hm
    from hitomi_amd import mocks

    n, L = mesh
    with hm.Context(0) as ctx:
        cat = mocks.survey_catalogue(ctx, hm.sky_patch(500.0, 800.0, 0.0, 40.0, -20.0, 20.0), 20_000, 100_000,
                                     seed=sum(n), n_side=64, randoms_on_device=False)
    hio.write_xyzw(tmp_path / "data.dat", cat["posD"], cat["wD"])
    hio.write_xyzw(tmp_path / "rand.dat", cat["posR"], cat["wR"])
    out = {}
    for prec in (None, "FP64"):
        ini = tmp_path / f"pk_{prec}.ini"
        hio.write_ini(ini, data_dir=str(tmp_path), data_file="data.dat", random_file="rand.dat",
                      output_dir=str(tmp_path / f"out_{prec}"), measure="pk", boxsize_x=L[0], boxsize_y=L[1],
                      boxsize_z=L[2], n_mesh_x=n[0], n_mesh_y=n[1], n_mesh_z=n[2], kmin=0.01, kmax=0.2, n_kbin=20,
                      precision=prec)
        r = run_survey(ini, tmp_path)
        assert r.returncode == 0, r.stdout
        out[prec] = hio.read_table(tmp_path / f"out_{prec}" / "pk0.dat")
    assert np.all(np.abs(out[None][:, 1] - out["FP64"][:, 1]) <= 1e-5 * np.abs(out["FP64"][:, 1]))
