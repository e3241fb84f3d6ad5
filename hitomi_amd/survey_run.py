#!/usr/bin/env python3
"""Survey-scale P_l timing: calcPowerSpectrum (powerspec.hpp:4-91) on a
512^3 mesh with the survey-geometry workload, one plan per multipole as the
reference runs one job per ELL.

Catalogue (hitomi_amd.mocks.survey_catalogue, seeded): 10^6 Zel'dovich data
in the shell 1200-1800 Mpc/h times RA 0-60 deg, Dec -30..30 deg, displaced
from a Gaussian field with P(k) = 2e4 (k/0.05)^-1.5 exp(-(k/0.5)^2); 5x10^7
randoms uniform in the same patch (generated on the device); FKP weights
1/(1 + nbar 10^4).  The FFT box is L = 2000 Mpc/h, 20 k-bins 0.015-0.205
h/Mpc, TSC.  Positions and weights are resident in HBM before the timed
executions.

Prints one JSON line (per-measurement times, a per-kernel breakdown of one
FP32 P_0 measurement from HIP events); --md writes a short summary.
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
from hitomi_amd import mocks  # noqa: E402

KERNELS = ["survey_g", "survey_wsums", "survey_minpos", "survey_shift", "survey_weights", "survey_sqsum", "survey_to_complex",
           "survey_sn", "survey_reduce", "survey_fft_z", "survey_fft_y", "survey_fft_x", "assign_count", "assign_scan", "assign_bucket", "assign_scatter", "assign_gather",
           "assign_border", "fill_zero", "fft_c2c_z", "fft_c2c_y", "fft_c2c_x", "bin_survey", "bin_reduce"]


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

    ctx = hm.Context(0)
    patch = hm.sky_patch(1200.0, 1800.0, 0.0, 60.0, -30.0, 30.0)
    cat = mocks.survey_catalogue(ctx, patch, int(args.nd), int(args.nr), seed=2024)
    dpd, dwd = ctx.array(cat["posD"]), ctx.array(cat["wD"])
    dpr, dwr = cat["posR"], cat["wR"]
    ctx.sync()
    kw = dict(n_kbin=20, kmin=0.015, kmax=0.205)
    out = {"workload": f"survey P_l, {args.mesh}^3 TSC, L = {args.boxsize:g}, {int(args.nd)} Zel'dovich data + "
                       f"{int(args.nr)} randoms in the shell 1200-1800 Mpc/h x RA 0-60 x Dec -30..30 deg, FKP "
                       "weights 1/(1 + nbar 1e4), 20 k-bins 0.015-0.205, one plan per ELL",
           "ms_per_measurement": {}, "first_bins": {}}
    for prec, name in ((hm.FP32, "fp32"), (hm.FP64, "fp64")):
        for ell in args.ells:
            plan = hm.SurveyPlan(ctx, args.mesh, args.boxsize, "TSC", ell=ell, precision=prec, **kw)
            pk, nm, info = plan.execute(dpd, dwd, dpr, dwr)  # warm-up, lazy allocations
            ctx.sync()
            t = time.perf_counter()
            for _ in range(args.reps):
                pk, nm, info = plan.execute(dpd, dwd, dpr, dwr)
            ctx.sync()
            ms = (time.perf_counter() - t) * 1000.0 / args.reps
            out["ms_per_measurement"][f"{name} P_{ell}"] = round(ms, 2)
            out["first_bins"][f"{name} P_{ell}"] = [float(f"{v:.6e}") for v in pk.real[:3]]
            if prec == hm.FP32 and ell in (0, 2):
                ctx.timing(True)
                ctx.timing_reset()
                plan.execute(dpd, dwd, dpr, dwr)
                ctx.sync()
                br = {}
                for k in KERNELS:
                    kms, nl = ctx.kernel_time(k)
                    if nl:
                        br[k] = [round(kms, 3), nl]
                ctx.timing(False)
                out[f"fp32_P{ell}_kernels_ms_launches"] = br
                out[f"fp32_P{ell}_kernel_total_ms"] = round(sum(v[0] for v in br.values()), 3)
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
            f.write("# Survey P_l at 512^3 with 1e6 data and 5e7 randoms\n\nCommand (repository root on the GPU host):"
                    "\n\n    python3 hitomi_amd/survey_run.py --md <file>\n\n```\n" + json.dumps(out, indent=1) +
                    "\n```\n")
    for a in (dpd, dwd, dpr, dwr):
        a.free()
    ctx.close()


if __name__ == "__main__":
    main()
