#!/usr/bin/env python3
"""Survey-scale P_l timing: calcPowerSpectrum (powerspec.hpp:4-91) on a
512^3 mesh with 10^6 data and 5x10^7 random points, one plan per multipole as
the reference runs one job per ELL.

Catalogue (synthetic, seeded): randoms uniform in the shell 600 < |x| < 1800
Mpc/h of the octant x, y, z > 0 (so lines of sight vary across the volume);
data drawn from the same volume with a density modulated along a plane wave,
weights 1 / (1 + 0.5 u), u uniform in [0, 1).  The FFT box is L = 2000 Mpc/h,
20 k-bins over 0.01-0.21 h/Mpc, TSC.  Positions and weights are resident in
HBM before the timed executions.

Prints one JSON line; --md writes a short summary.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import hitomi_amd as hm  # noqa: E402


def shell_points(rng, n, rmin=600.0, rmax=1800.0):
    """n points uniform in the octant shell rmin < |x| < rmax, x, y, z > 0."""
    r = np.cbrt(rng.uniform(rmin ** 3, rmax ** 3, n))
    mu = rng.uniform(0.0, 1.0, n)
    phi = rng.uniform(0.0, 0.5 * np.pi, n)
    s = np.sqrt(1.0 - mu * mu)
    return np.stack([r * s * np.cos(phi), r * s * np.sin(phi), r * mu], 1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mesh", type=int, default=512)
    ap.add_argument("--boxsize", type=float, default=2000.0)
    ap.add_argument("--nd", type=float, default=1e6)
    ap.add_argument("--nr", type=float, default=5e7)
    ap.add_argument("--ells", type=int, nargs="+", default=[0, 2, 4])
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--md", default=None)
    args = ap.parse_args()

    rng = np.random.default_rng(2024)
    nd, nr = int(args.nd), int(args.nr)
    pr = shell_points(rng, nr)
    wr = 1.0 / (1.0 + 0.5 * rng.random(nr))
    # data: thin a larger draw by 1 + 0.5 cos(k.x) so that the field has a signal
    cand = shell_points(rng, 3 * nd)
    keep = rng.random(len(cand)) < (1.0 + 0.5 * np.cos(cand @ np.array([0.02, 0.015, 0.01]))) / 1.5
    pd = cand[keep][:nd]
    wd = 1.0 / (1.0 + 0.5 * rng.random(len(pd)))

    ctx = hm.Context(0)
    dpd, dwd = ctx.array(np.ascontiguousarray(pd)), ctx.array(wd)
    dpr, dwr = ctx.array(np.ascontiguousarray(pr)), ctx.array(wr)
    ctx.sync()
    out = {"workload": f"survey P_l, {args.mesh}^3 TSC, L = {args.boxsize:g}, {len(pd)} data + {nr} randoms "
                       "(octant shell 600-1800 Mpc/h), 20 k-bins 0.01-0.21, one plan per ELL",
           "ms_per_measurement": {}, "first_bins": {}}
    for prec, name in ((hm.FP32, "fp32"), (hm.FP64, "fp64")):
        for ell in args.ells:
            plan = hm.SurveyPlan(ctx, args.mesh, args.boxsize, "TSC", n_kbin=20, kmin=0.01, kmax=0.21, ell=ell,
                                 precision=prec)
            pk, nm, info = plan.execute(dpd, dwd, dpr, dwr)  # warm-up, lazy allocations
            ctx.sync()
            t = time.perf_counter()
            for _ in range(args.reps):
                pk, nm, info = plan.execute(dpd, dwd, dpr, dwr)
            ctx.sync()
            ms = (time.perf_counter() - t) * 1000.0 / args.reps
            out["ms_per_measurement"][f"{name} P_{ell}"] = round(ms, 2)
            out["first_bins"][f"{name} P_{ell}"] = [float(f"{v:.6e}") for v in pk.real[:3]]
            plan.close()
            print(f"{name} P_{ell}: {ms:.2f} ms", file=sys.stderr, flush=True)
    p32 = np.array(out["first_bins"]["fp32 P_0"])
    p64 = np.array(out["first_bins"]["fp64 P_0"])
    out["fp32_vs_fp64_P0_first_bins_rel"] = float(np.abs(p32 - p64).max() / np.abs(p64).max())
    out["alpha"] = info["alpha"]
    out["vsurvey"] = info["vsurvey"]
    print(json.dumps(out), flush=True)
    if args.md:
        with open(args.md, "w") as f:
            f.write("# Survey P_l at 512^3 with 5e7 randoms\n\nCommand (repository root on the GPU host):\n\n"
                    "    python3 hitomi_amd/survey_run.py --md <file>\n\n```\n" + json.dumps(out, indent=1) + "\n```\n")
    for a in (dpd, dwd, dpr, dwr):
        a.free()
    ctx.close()


if __name__ == "__main__":
    main()
