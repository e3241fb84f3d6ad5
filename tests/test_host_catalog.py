"""The periodic-box program's catalogue readers on the CPU (host code of the
drop-in, hitomi_amd/host/hm_catalog.hpp), against independent restatements
of the reference's rules:
  - Rockstar halo lists (particleGadget.hpp:171-209, 300-388): cosmology from
    the '#a =', '#Om = ...; Ol = ...; h = ...', '#Box size:' and
    '#Particle mass:' header lines; rows of exactly 34 columns; halos with
    log10_Mmin < log10(M) < log10_Mmax (column 20); position and velocity
    from columns 9-14;
  - setRSD (particleGadget.hpp:533-563): z += v_z / (a H(a)),
    H = 100 sqrt(Om a^-3 + (1 - Om - Ol) a^-2 + Ol), then one periodic shift
    per axis (particleGadget.hpp:565-591);
  - Gadget-1 velocities times sqrt(a) before the shift (particleGadget.hpp:271).
"""
import os
import subprocess

import numpy as np
import pytest

from hitomi_amd import io as hio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = tmp_path_factory.mktemp("probe") / "catalog_probe"
    src = os.path.join(ROOT, "tests", "catalog_probe.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), src], check=True)
    return str(exe)


def run_probe(probe, *args):
    r = subprocess.run([probe, *map(str, args)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    head = lines[0].split()
    rows = np.array([[float(v) for v in ln.split()] for ln in lines[1:]]).reshape(-1, 6)
    return int(head[0]), [float(v) for v in head[1:]], rows


def hubble(a, om, ol):
    return 100.0 * np.sqrt(om / a ** 3 + (1.0 - om - ol) / a ** 2 + ol)


def write_rockstar(path, rows, a=0.5, om=0.307115, ol=0.692885, h=0.6777, box=1000.0, mass=1.40068e9):
    with open(path, "w") as f:
        f.write("#ID DescID Mvir Vmax Vrms Rvir Rs Np X Y Z VX VY VZ JX JY JZ Spin rs_klypin Mvir_all M200b M200c "
                "M500c M2500c Xoff Voff spin_bullock b_to_a c_to_a A[x] A[y] A[z] T/|U| PID\n")
        f.write(f"#a = {a:.6f}\n")
        f.write(f"#Om = {om:.6f}; Ol = {ol:.6f}; h = {h:.6f}\n")
        f.write("#FOF linking length: 0.280000\n")
        f.write(f"#Particle mass: {mass:.5e} Msun/h\n")
        f.write(f"#Box size: {box:.6f} Mpc/h\n")
        for r in rows:
            f.write(r + "\n")


def rockstar_rows(n, seed, box):
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(n):
        c = rng.random(32) * 10.0
        c[7:10] = rng.random(3) * box                    # X Y Z (columns 9-11)
        c[10:13] = rng.normal(0.0, 300.0, 3)             # VX VY VZ (columns 12-14)
        c[18] = 10.0 ** rng.uniform(11.5, 14.5)          # column 20: the mass the reference cuts on
        rows.append(f"{i} " + " ".join(f"{v:.6e}" for v in c) + f" {-1 if i % 3 else i - 1}")
        cols.append(c)
    return rows, np.array(cols)


def test_rockstar_reader_mass_cut_and_columns(probe, tmp_path):
    rows, cols = rockstar_rows(400, 1, 1000.0)
    rows.insert(5, "17 1.0 2.0 3.0")                      # too few columns: skipped
    rows.insert(9, "# a comment line")
    f = tmp_path / "halos.list"
    write_rockstar(f, rows)
    n, (a, om, ol, h, box), got = run_probe(probe, "rockstar", f, 12.9, 13.5, 0)
    # the header values as written (%.6f)
    assert (a, om, ol, h, box) == (0.5, 0.307115, 0.692885, 0.6777, 1000.0)
    parsed = np.array([[float(f"{v:.6e}") for v in c] for c in cols])
    lm = np.log10(parsed[:, 18])
    keep = (12.9 < lm) & (lm < 13.5)
    assert n == keep.sum() and 0 < n < len(cols)
    assert np.array_equal(got[:, :3], parsed[keep][:, 7:10])
    assert np.array_equal(got[:, 3:], parsed[keep][:, 10:13])


def test_rockstar_rsd_shift_and_wrap(probe, tmp_path):
    rows, cols = rockstar_rows(300, 2, 500.0)
    f = tmp_path / "halos.list"
    write_rockstar(f, rows, a=0.4, om=0.3, ol=0.7, box=500.0)
    n, _, plain = run_probe(probe, "rockstar", f, 11.0, 15.0, 0)
    _, _, shifted = run_probe(probe, "rockstar", f, 11.0, 15.0, 1)
    z = plain[:, 2] + plain[:, 5] / (0.4 * hubble(0.4, 0.3, 0.7))
    z = np.where(z >= 500.0, z - 500.0, np.where(z < 0.0, z + 500.0, z))
    assert np.abs(shifted[:, 2] - z).max() < 1e-9
    assert np.array_equal(shifted[:, :2], plain[:, :2])
    moved = np.abs(plain[:, 5]) > 0
    assert moved.any() and (np.abs(shifted[:, 2] - plain[:, 2]) > 0).sum() == moved.sum()
    assert ((shifted[:, 2] >= 0.0) & (shifted[:, 2] < 500.0)).all()


def test_gadget_velocities_and_rsd(probe, tmp_path):
    rng = np.random.default_rng(3)
    L, a = 1000.0, 0.5
    pos = (rng.random((5000, 3)) * L).astype(np.float32)
    vel = rng.normal(0.0, 500.0, (5000, 3)).astype(np.float32)
    hio.write_gadget1(str(tmp_path / "snap"), pos, vel, boxsize=L, a=a, omega_m=0.3, omega_l=0.7, num_files=2)
    n, (aa, om, ol, _, box), plain = run_probe(probe, "gadget", tmp_path / "snap", 0)
    assert n == 5000 and aa == a and box == L
    assert np.array_equal(plain[:, :3], pos.astype(np.float64))
    assert np.abs(plain[:, 3:] - vel.astype(np.float64) * np.sqrt(a)).max() == 0.0
    _, _, shifted = run_probe(probe, "gadget", tmp_path / "snap", 1)
    z = plain[:, 2] + plain[:, 5] / (a * hubble(a, om, ol))
    z = np.where(z >= L, z - L, np.where(z < 0.0, z + L, z))
    assert np.abs(shifted[:, 2] - z).max() < 1e-9
