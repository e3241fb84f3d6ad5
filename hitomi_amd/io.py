"""Catalogue and result file formats of the reference, for writing inputs to
and reading outputs of the drop-in programs.

ASCII catalogues `x y z w` in %.10e (write_xyzw).  Gadget-1 snapshot files `<base>.<i>` (particleGadget.hpp:35-51, 211-298): a
256-byte header between 4-byte record markers, then float32 positions,
float32 velocities (the reader multiplies them by sqrt(a)) and int64 ids,
each block between record markers.  Result files pk%d.dat / bk%d%d%d_%02d.dat
are tab-separated text with one '##' header line (powerspec.hpp:321-326,
bispec.hpp:1191-1195).
"""
from __future__ import annotations

import struct

import numpy as np

HEADER_FMT = "<6I6dddii6Iii4d"


def write_gadget1(base, pos, vel=None, boxsize=1000.0, a=1.0, omega_m=0.3, omega_l=0.7, h=0.7, num_files=1,
                  mass=1.0):
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
    n = pos.shape[0]
    vel = np.zeros_like(pos) if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(-1, 3)
    ids = np.arange(n, dtype=np.int64)
    splits = np.array_split(np.arange(n), num_files)
    for i, idx in enumerate(splits):
        npart = [0, len(idx), 0, 0, 0, 0]
        ntot = [0, n, 0, 0, 0, 0]
        hdr = struct.pack(HEADER_FMT, *npart, 0.0, mass, 0.0, 0.0, 0.0, 0.0, a, 1.0 / a - 1.0, 0, 0, *ntot, 0,
                          num_files, boxsize, omega_m, omega_l, h)
        hdr += b"\0" * (256 - len(hdr))
        assert len(hdr) == 256
        with open(f"{base}.{i}", "wb") as f:
            for block in (hdr, pos[idx].tobytes(), vel[idx].tobytes(), ids[idx].tobytes()):
                f.write(struct.pack("<i", len(block)))
                f.write(block)
                f.write(struct.pack("<i", len(block)))


def write_xyzw(path, pos, w):
    """A catalogue as the reference's converters write it (np.savetxt(..., fmt="%.10e") of the columns
    x, y, z, w: fits2dat_galaxy_weight.py:134, 158), the input of ParticleBOSSClass (particle.hpp:93-154)."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    w = np.asarray(w, dtype=np.float64).reshape(-1)
    np.savetxt(path, np.column_stack([pos, w]), fmt="%.10e")


def read_table(path):
    """Numeric rows of a result file (the '##' header line skipped)."""
    rows = []
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            rows.append([float(x) for x in line.split()])
    return np.array(rows)


def write_ini(path, **kw):
    """A parameter file in the reference's `key = value %% comment` form."""
    defaults = dict(data_dir=".", data_file="snap", random_file="AAA", output_dir="results", measure="pk",
                    realization=0, ell1=0, ell2=0, ELL=0, boxsize_x=1000.0, boxsize_y=1000.0, boxsize_z=1000.0,
                    n_mesh_x=64, n_mesh_y=64, n_mesh_z=64, kmin=0.01, kmax=0.2, n_kbin=20, rmin=30.0, rmax=150.0,
                    n_rbin=25, flag_recon="False", b1_fid=1.0, RG=15.0, n_mesh_recon_x=64, n_mesh_recon_y=64,
                    n_mesh_recon_z=64, ith_kbin=0, ith_rbin=0, NR=0, data_mpi_file="mock", sim_data="Gadget",
                    flag_RSD="False", log10_Mmin=12.9, log10_Mmax=13.5, assignment="TSC")
    defaults.update(kw)
    with open(path, "w") as f:
        f.write("%%%%%%%%%%\n%% periodic-box measurement\n%%%%%%%%%%\n")
        for k, v in defaults.items():
            if v is None:
                continue
            f.write(f"{k} = {v} %% {k}\n")
