#!/usr/bin/env python3
"""bench.py -- mocks/hour of the periodic-box estimator on one MI355X.

One step = one mock through the whole measurement the reference spreads over
3 + 20 jobs (main_box.cpp with measure = pk for l = 0, 2, 4 and measure = bk
for every ith_kbin): P_0, P_2, P_4 and all 210 B_000(k1, k2) pairs with their
shot noise, on a 512^3 TSC mesh, L = 2000 Mpc/h, 4,000,000 Zel'dovich
particles (a 256^3 lattice, seeded subset) from a seeded Gaussian field with
P(k) = 2e4 (k/0.05)^-1.5 exp(-(k/0.5)^2) [1 + 0.05 sin(2 pi k / 0.0628)],
20 k-bins with centres 0.015 .. 0.205 h/Mpc.  The mocks are generated on the
device and resident in HBM before the timed region; the steps run through
the batched pipeline (each mock's data vector formed on the device, no host
synchronisation between mocks).

The covariance run (--cov-mocks, default 512) is the batched covariance
workload: mocks of uneven size (N ~ Poisson(lambda), lambda ~ Gamma with
mean 1.5e6 and 25% scatter, truncated to [0.8e6, 2.5e6]), L = 1500 Mpc/h, 256^3 mesh, each generated on the device
with its own seed and measured (P_0, P_2, P_4 and B_000) through the
graph-replayed pipeline, ending with the 60 x 60 FP64 covariance of
[P_0, P_2, P_4] x 20 bins.

Prints one JSON line (see the repository's bench contract).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "mocks/hour (P_ℓ+B_000, 512³) batched FP32 vs plain FP64 GPU; HBM GB/s"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector, counting packed (two-wide) FMA

N_MESH, BOX, N_SIDE, N_PART = 512, 2000.0, 256, 4_000_000
KW = dict(n_kbin=20, kmin=0.015, kmax=0.205, pk_ells=(0, 2, 4), bk=True)
PK = dict(amp=2.0e4, k0=0.05, slope=-1.5, kcut=0.5, bao_amp=0.05, bao_period=0.0628)
COV_MESH, COV_BOX, COV_SIDE = 256, 1500.0, 256


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def support(n, L, kmin, kmax, nk):
    """Largest |n_a| of any mode that can land in a bin (as the plan computes it)."""
    kb = kmin + (kmax - kmin) / (nk - 1) * np.arange(nk)
    dkb = kb[1] - kb[0]
    umax = float(np.max(kb + dkb * 0.5))
    s = int(np.floor((umax + 1e-4) / (2 * np.pi / L))) + 1
    return -1 if (2 * s + 1 >= n or s >= n // 2) else s


def shell_widths(L, kmin, kmax, nk, s):
    """Per-shell cube widths 2 s_j + 1 on the triple mesh: every mode of shell j
    has |k| <= upper_j, so |n_a| <= upper_j / dk (+1 for rounding), capped at s."""
    kb = kmin + (kmax - kmin) / (nk - 1) * np.arange(nk)
    dkb = kb[1] - kb[0]
    dk = 2 * np.pi / L
    return [2 * min(int(np.floor((k + 0.5 * dkb) / dk)) + 1, s) + 1 for k in kb]


def n_groups(n):
    """Distinct i'^2 + j'^2 over minimum-image row indices: the rows the shot-noise xi keeps."""
    h = np.minimum(np.arange(n), n - np.arange(n)) ** 2
    return int(np.unique(h[:, None] + h[None, :]).size)


def algorithmic_bytes(name, n, M, s, nk, npart, elem=8, ws=None):
    """Compulsory HBM bytes of one launch of `name` (FP32 complex = 8 B);
    ws: the shells' own cube widths (default: all 2 s + 1)."""
    half = n * n * (n // 2 + 1) * elem
    w = 2 * s + 1
    ws = ws or [w] * nk
    W3 = 4 * s + 1 if 4 * s + 1 < M else M
    table = {
        "assign_gather": npart * 24 + n ** 3 * 8,                # sorted positions in, int64 mesh out
        "xi_group": half,                                         # the y-transformed half spectrum once
        "fft_r2c_z": n ** 3 * 8 + half,               # int64 mesh in, half spectrum out
        "fft_r2c_y": 2 * half,
        "fft_r2c_x": 2 * half,
        "xi_pass_x": 2 * half,
        "fft_x_xi": 3 * half,                         # the spectrum in once; it and xi's x pass out
        "xi_planes_x": 2 * 2 * n * n * elem,          # the planes k_z = 0, n_z/2 of xi's x pass
        "xi_pass_y": 2 * half,
        "xi_pass_z": n_groups(n) * ((n // 2 + 1) * elem + n * (elem // 2)),  # grouped rows, c2r
        # triple-product fields: shell j on its own cube of width ws[j], one delta/W of support W3
        # (compulsory in + out)
        "tri_stage": W3 ** 3 * elem * 2 + w ** 3 * 2,
        "tri_pass_x": sum(v ** 3 * (elem + 2) + M * v * v * elem for v in ws),
        "tri3_pass_x": (W3 ** 3 + M * W3 * W3) * elem,
        "tri_pass_y": sum((M * v * v + M * M * v) * elem for v in ws),
        "tri3_pass_y": (M * W3 * W3 + M * M * W3) * elem,
        "tri_z_gram": (sum(M * M * v for v in ws) + M * M * W3) * elem,
        "gram": (nk + 1) * M ** 3 * elem,
    }
    return table.get(name)


def tri_z_gram_flops(M, nk):
    """Arithmetic of one tri_z_gram launch over the M^2 rows of the triple
    mesh: per row the z transforms of the nk shell lines and the delta/W line
    (5 M log2 M each), the products F_a delta/W (6 M per shell) and the
    nk (nk + 1) / 2 complex pair sums (8 M each)."""
    per_row = (nk + 1) * 5.0 * M * np.log2(M) + nk * 6.0 * M + nk * (nk + 1) / 2 * 8.0 * M
    return M * M * per_row


# HBM bytes per launch of the dominant kernels, measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE on this workload and scaled by the counters' reading
# of a 1 GiB copy (x2.000 for FETCH_SIZE, x1.000 for WRITE_SIZE).
MEASURED_TRAFFIC = {"tri_z_gram": 1245.0e6}


def make_mocks(ctx, count, seed0=1000):
    import hitomi_amd as hm

    cfg = hm.mock_config(N_SIDE, BOX, **PK)
    out = []
    for i in range(count):
        t = time.time()
        out.append(hm.mock_zeldovich(ctx, cfg, seed0 + i, N_PART))
        ctx.sync()
        log(f"mock {i}: {N_PART} particles generated on the device in {time.time() - t:.2f}s")
    return out


def covariance_run(ctx, n_mocks, lanes, seed0=5000):
    """The batched covariance workload; returns its summary."""
    import hitomi_amd as hm
    from hitomi_amd import mocks

    sizes = mocks.poisson_sizes(n_mocks, 1_500_000, 800_000, 2_500_000, seed=seed0, scatter=0.25)
    cfg = hm.mock_config(COV_SIDE, COV_BOX, **PK)
    buf = ctx.empty((int(sizes.max()), 3), np.float64)
    # the generator alone, for its share of the time
    hm.mock_zeldovich(ctx, cfg, seed0, int(sizes[0]), out=buf)
    ctx.sync()
    t = time.perf_counter()
    for i in range(8):
        hm.mock_zeldovich(ctx, cfg, seed0 + i, int(sizes[i]), out=buf)
    ctx.sync()
    gen_ms = (time.perf_counter() - t) * 1000.0 / 8
    out = {"mocks": int(n_mocks), "lanes": int(lanes), "n_mesh": COV_MESH, "boxsize": COV_BOX, "lattice": COV_SIDE,
           "particles_min": int(sizes.min()), "particles_max": int(sizes.max()),
           "particles_mean": float(sizes.mean()), "generator_ms_per_mock": round(gen_ms, 3)}
    for graph in (True, False):
        b = hm.Batch(ctx, COV_MESH, COV_BOX, max_mocks=0, use_graph=graph, lanes=lanes, precision=hm.FP32, **KW)
        hm.mock_zeldovich(ctx, cfg, seed0, int(sizes[0]), out=buf)
        out[f"{'graph' if graph else 'eager'}_autotune"] = b.autotune(buf, reps=3, npart=int(sizes[0])).split(
            "chosen: ")[-1]
        b.push(buf, int(sizes[0]))  # warm-up (and, with the graph, the capture on the next push)
        for w in range(1, 2 * lanes):
            hm.mock_zeldovich(ctx, cfg, seed0 + w, int(sizes[w]), out=buf)
            b.push(buf, int(sizes[w]))
        b.reset()
        ctx.sync()
        t = time.perf_counter()
        for i, n in enumerate(sizes):
            hm.mock_zeldovich(ctx, cfg, seed0 + i, int(n), out=buf)
            b.push(buf, int(n))
        mean, cov = b.covariance()
        ms = (time.perf_counter() - t) * 1000.0 / n_mocks
        key = "graph" if graph else "eager"
        out[f"{key}_ms_per_mock"] = round(ms, 3)
        if graph:
            out["mocks_per_hour"] = round(3600e3 / ms, 1)
            out["vector"] = f"{b.vec_len} entries ([P_0, P_2, P_4] x 20, then B_000 on 210 bin pairs)"
            out["covariance"] = f"{b.cov_len} x {b.cov_len} FP64 ([P_0, P_2, P_4] x 20), N - 1 estimator"
            sd = np.sqrt(np.diag(cov))
            out["p0_rel_sigma_k1"] = round(float(sd[0] / abs(mean[0])), 5)
            out["finite"] = bool(np.isfinite(cov).all())
        b.close()
        log(f"covariance run ({key}): {ms:.3f} ms per mock incl. generation ({gen_ms:.3f} ms)")
    buf.free()
    return out


def host_facts():
    """nproc and the CPU model of the host the baseline runs on."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_baseline(pos):
    """The FP64 CPU oracle (serial, 1 core) on a bounded sample of the same
    workload: one P_0 job of the reference's job structure at the full 512^3
    size on the same catalogue (assignment, FFT, binning), and one B_000 job
    (ith_kbin = 0, 20 k2 bins) at 128^3, whose mesh-proportional part is scaled
    x64 to 512^3 (this ignores the extra log N of the FFTs and so favours the
    CPU); one mock = 3 P jobs + 20 B jobs."""
    from oracle import pyoracle as po

    p512 = po.param(N_MESH, BOX, po.TSC, kmin=KW["kmin"], kmax=KW["kmax"], n_kbin=KW["n_kbin"])
    t = time.time()
    po.box_pk(p512, pos)
    t_pk = time.time() - t
    n = 128
    p = po.param(n, BOX, po.TSC, kmin=KW["kmin"], kmax=KW["kmax"], n_kbin=KW["n_kbin"])
    t = time.time()
    po.box_density_fluctuation(p, pos)
    t_assign = time.time() - t
    t = time.time()
    pb = po.param(n, BOX, po.TSC, kmin=KW["kmin"], kmax=KW["kmax"], n_kbin=KW["n_kbin"], ith_kbin=0)
    po.box_bk(pb, pos)
    t_bk = time.time() - t
    scale = (N_MESH / n) ** 3
    bk_job = 6 * t_assign + scale * max(t_bk - 6 * t_assign, 0.0)
    sec_per_mock = 3 * t_pk + 20 * bk_job
    sample = (f"oracle FP64, 1 core: P_0 job at 512^3 {t_pk:.1f}s (timed at full size), B_000 job (ith_kbin=0) "
              f"{t_bk:.2f}s at 128^3 (assignment {t_assign:.2f}s) scaled x{scale:.0f} (mesh part) to 512^3; "
              f"3 P + 20 B jobs = {sec_per_mock:.0f} s/mock, on the 4.0e6-particle mock")
    return {"value": 3600.0 / sec_per_mock, "unit": "mocks/hour", "cores": 1, "kind": "port", "sample": sample,
            "sample_seconds": round(t_pk + t_assign + t_bk, 1), "p0_job_512_s": round(t_pk, 2),
            "bk_job_512_s_scaled": round(bk_job, 1), "host": host_facts()}


def copy_calibration(ctx, nbytes):
    """One device-to-device copy of nbytes with the library's copy kernel
    (hm::k_copy16): the known-size kernel that scales the FETCH_SIZE /
    WRITE_SIZE counters."""
    a = ctx.empty((nbytes // 16, 2), np.float64)
    b = ctx.empty((nbytes // 16, 2), np.float64)
    ctx.check(ctx.lib.hm_memset_zero(ctx.h, a.ptr, a.nbytes), "memset")
    ctx.sync()
    b.copy_from(a)
    ctx.sync()
    log(f"copy calibration: {a.nbytes} bytes read, {a.nbytes} bytes written")
    a.free()
    b.free()


def run(args):
    import hitomi_amd as hm

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: collectives, sharding and multi-GPU runs are out "
                                   "of scope; run with --gpus 1"}))
        sys.exit(2)
    ctx = hm.Context(0)
    dpos = make_mocks(ctx, args.mocks)
    npart = N_PART
    plan = hm.BoxPlan(ctx, N_MESH, BOX, "TSC", precision=hm.FP32, **KW)
    M = plan.triple_mesh[0]
    s = support(N_MESH, BOX, KW["kmin"], KW["kmax"], KW["n_kbin"])
    ws = shell_widths(BOX, KW["kmin"], KW["kmax"], KW["n_kbin"], s)
    log(f"plan: triple mesh {plan.triple_mesh}, support {s}, device bytes {plan.device_bytes / 2**30:.2f} GiB")

    # per-kernel breakdown of one (untimed) mock
    plan.execute(dpos[0])
    ctx.timing(True)
    ctx.timing_select(None)
    ctx.timing_reset()
    plan.execute(dpos[0])
    rows = []
    for name in hm.BOX_KERNELS:
        ms, nl = ctx.kernel_time(name)
        if nl:
            rows.append((ms, name, nl))
    rows.sort(reverse=True)
    tot = sum(r[0] for r in rows)
    log(f"kernel breakdown of one mock ({tot:.2f} ms of kernel time):")
    for ms, name, nl in rows:
        b = algorithmic_bytes(name, N_MESH, M, s, KW["n_kbin"], npart, ws=ws)
        bw = f"{b / (ms / nl) / 1e6:8.0f} GB/s" if b else ""
        log(f"  {name:14s} {ms:8.3f} ms  {nl:3d} launches  {bw}")
    dom = next(r for r in rows if algorithmic_bytes(r[1], N_MESH, M, s, KW["n_kbin"], npart, ws=ws))
    dom_name = dom[1]
    ctx.timing(False)
    r = plan.execute(dpos[0])
    ref_vec = np.concatenate([r["pk"][e, :, 0] for e in range(3)] + [r["bk"][..., 0][np.triu_indices(KW["n_kbin"])]])
    plan.close()

    # timed region: the batched pipeline (eager launches, so that HIP events can
    # bracket the dominant kernel's launches), its launch shapes autotuned first
    batch = hm.Batch(ctx, N_MESH, BOX, max_mocks=1, use_graph=False, lanes=args.lanes, precision=hm.FP32, **KW)
    tune = batch.autotune(dpos[0], reps=2) if args.autotune else "off"
    log(f"autotune: {tune}")
    for i in range(max(args.warmup, args.lanes)):
        batch.push(dpos[i % len(dpos)])
    batch.reset()
    ctx.timing(True)
    ctx.timing_select(dom_name)
    ctx.timing_reset()
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        batch.push(dpos[i % len(dpos)])
    batch.collect_timing()  # waits for every lane
    t1 = time.perf_counter()
    ms_dom, n_dom = ctx.kernel_time(dom_name)
    busy_dom = ctx.kernel_busy(dom_name)
    ctx.timing(False)
    same = bool(np.array_equal(batch.vectors()[0], ref_vec))
    batch.close()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    value = 3600.0 * 1000.0 / ms_per_step
    b_dom = algorithmic_bytes(dom_name, N_MESH, M, s, KW["n_kbin"], npart, ws=ws)
    # the lanes' launches overlap: per launch, the kernel's busy time (the union
    # of its launch intervals) divided by the launch count
    ms_launch = busy_dom / n_dom
    achieved = b_dom / ms_launch / 1e6
    fl_dom = tri_z_gram_flops(M, KW["n_kbin"]) if dom_name == "tri_z_gram" else None
    log(f"batched: {ms_per_step:.3f} ms per mock; first vector equals the per-mock plan's: {same}")

    # the same steps replayed from the captured graph
    gbatch = hm.Batch(ctx, N_MESH, BOX, max_mocks=1, use_graph=True, lanes=args.lanes, precision=hm.FP32, **KW)
    if args.autotune:
        gbatch.autotune(dpos[0], reps=2)
    for i in range(max(args.warmup, 2 * args.lanes)):
        gbatch.push(dpos[i % len(dpos)])
    gbatch.collect_timing()
    t = time.perf_counter()
    for i in range(args.steps):
        gbatch.push(dpos[i % len(dpos)])
    gbatch.collect_timing()
    ms_graph = (time.perf_counter() - t) * 1000.0 / args.steps
    gbatch.close()
    log(f"graph replay: {ms_graph:.3f} ms per mock")
    if args.copy_calibration:
        copy_calibration(ctx, args.copy_calibration)

    plain = None
    plain_regrid = None
    err = None
    if args.plain_steps > 0:
        pplan = hm.BoxPlan(ctx, N_MESH, BOX, "TSC", precision=hm.FP64, regrid=False, prune=False, **KW)
        r64 = pplan.execute(dpos[0])
        # the FP32 batched path against the plain FP64 one on the same catalogue, per multipole
        nk = KW["n_kbin"]
        p32 = ref_vec[:3 * nk].reshape(3, nk)
        p64 = r64["pk"][..., 0]
        b32 = ref_vec[3 * nk:]
        b64 = r64["bk"][..., 0][np.triu_indices(nk)]
        err = {f"P_{l}": float(np.abs(p32[e] - p64[e]).max() / np.abs(p64[0]).max()) for e, l in enumerate((0, 2, 4))}
        err["P_2 rel. where |P_2| > 0.1 |P_0|"] = float(np.max(np.where(np.abs(p64[1]) > 0.1 * np.abs(p64[0]),
                                                                     np.abs(p32[1] - p64[1]) / np.abs(p64[1]), 0.0)))
        err["B_000"] = float(np.abs(b32 - b64).max() / np.abs(b64).max())
        err = {k: float(f"{v:.3e}") for k, v in err.items()}
        err["normalisation"] = "max over k of |FP32 - FP64| / max|P_0 (FP64)|; B: / max|B (FP64)|"
        ctx.sync()
        t = time.perf_counter()
        for i in range(args.plain_steps):
            pplan.execute(dpos[i % len(dpos)])
        ctx.sync()
        ms_plain = (time.perf_counter() - t) * 1000.0 / args.plain_steps
        pplan.close()
        plain = {"mocks_per_hour": round(3600e3 / ms_plain, 1), "ms_per_mock": round(ms_plain, 2),
                 "what": "FP64 grids and FFTs, full 512^3 shell fields (no re-grid, no pruning), one mock at a time"}
        # the FP32 path's algorithm (exact re-grid, pruned passes) in FP64, one mock at a time: what the FP32
        # arithmetic alone earns
        fplan = hm.BoxPlan(ctx, N_MESH, BOX, "TSC", precision=hm.FP64, **KW)
        fplan.execute(dpos[0])
        ctx.sync()
        t = time.perf_counter()
        for i in range(args.plain_steps):
            fplan.execute(dpos[i % len(dpos)])
        ctx.sync()
        ms_f64 = (time.perf_counter() - t) * 1000.0 / args.plain_steps
        fplan.close()
        plain_regrid = {"ms_per_mock": round(ms_f64, 2), "mocks_per_hour": round(3600e3 / ms_f64, 1),
                        "what": "FP64 grids and FFTs with the FP32 path's algorithm (exact re-grid to the triple "
                                "mesh, pruned passes), one mock at a time (no batch lanes)"}

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "mocks/hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "periodic-box P_l (l=0,2,4) + B_000 all 210 (k1,k2) pairs with shot noise, one mock per step",
            "n_mesh": N_MESH, "boxsize": BOX, "assignment": "TSC", "particles": npart,
            "catalogue": "Zel'dovich, seeded 4,000,000-site subset of a 256^3 lattice, seeded Gaussian field with "
                         "BAO wiggle, generated on the device",
            "k_bins": "20 bins, centres 0.015-0.205 h/Mpc", "triple_mesh": M, "parallelism": "1 GPU",
            "pipeline": f"batched (device-side data vectors, no host synchronisation between mocks), {args.lanes} lanes "
                        "(plans on streams of their own taking mocks in turn)",
        },
        "batched_vector_equals_per_mock": same,
        "autotune": tune.split("chosen: ")[-1] if "chosen" in tune else tune,
        "graph_replay_ms_per_step": round(ms_graph, 3),
    }
    hbm = {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
           "traffic": MEASURED_TRAFFIC.get(dom_name),
           "algorithmic_bytes_per_launch": int(b_dom), "ms_per_launch": round(ms_launch, 4),
           "launches": n_dom, "busy_ms": round(busy_dom, 3),
           "avg_launch_ms": round(ms_dom / n_dom, 4),
           "avg_launch_ms_alone": round(dom[0] / dom[2], 4),
           "achieved_alone": round(b_dom / (dom[0] / dom[2]) / 1e6, 1),
           "frac_alone": round(b_dom / (dom[0] / dom[2]) / 1e6 / HBM_PEAK_GBS, 4),
           "note": "HIP events on each lane's stream around every launch in the timed steps; ms_per_launch "
                   "= busy_ms (the union of the launch intervals: overlapping launches of different lanes "
                   "count once) / launches; avg_launch_ms = the mean interval, which includes the slowdown "
                   "of sharing the GPU with the other lanes; *_alone: one launch with the GPU to itself"}
    if fl_dom:
        # the roof that binds: FP32 arithmetic (the SQ counters show the vector ALUs issuing most of the time
        # and HBM traffic at a few percent of its peak)
        out["roofline"] = {
            "bound": "mfma", "kernel": dom_name, "unit": "TFLOP/s", "flops_per_launch": int(fl_dom),
            "peak": FP32_PEAK_TFLOPS, "achieved": round(fl_dom / ms_launch / 1e9, 2),
            "frac": round(fl_dom / ms_launch / 1e9 / FP32_PEAK_TFLOPS, 4),
            "traffic": MEASURED_TRAFFIC.get(dom_name), "ms_per_launch": round(ms_launch, 4), "launches": n_dom,
            "achieved_alone": round(fl_dom / (dom[0] / dom[2]) / 1e9, 2),
            "frac_alone": round(fl_dom / (dom[0] / dom[2]) / 1e9 / FP32_PEAK_TFLOPS, 4),
            "note": "compute-bound on FP32 arithmetic: the z transforms of 21 lines per row in LDS and the 210 "
                    "complex pair sums run on the vector ALUs (the SQ counters show them issuing most of the "
                    "time; moving the pair sums to v_mfma_f32_16x16x4_f32 did not help because the LDS "
                    "transforms are the larger part); bound 'mfma' stands for the compute roof; peak = MI355X's "
                    "published FP32 rate; the code is built without packed FP32, which halves the vector rate it "
                    "can reach (78.65 TFLOP/s); traffic = measured HBM bytes per launch; the HBM view of the same "
                    "launches is roofline_hbm",
        }
        out["roofline_hbm"] = hbm
    else:
        out["roofline"] = hbm
    if plain:
        out["plain_fp64"] = plain
        out["fp64_regrid_prune"] = plain_regrid
        out["speedup_vs_fp64_regrid_prune"] = round(value / plain_regrid["mocks_per_hour"], 2)
        out["fp32_error_vs_fp64"] = err
        out["speedup_vs_plain_fp64"] = round(value / plain["mocks_per_hour"], 2)
    if args.cov_mocks > 0:
        out["covariance_run"] = covariance_run(ctx, args.cov_mocks, args.lanes)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(dpos[0].download())
    ctx.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mocks", type=int, default=2, help="distinct resident mocks cycled through")
    ap.add_argument("--plain-steps", type=int, default=2, help="steps of the plain FP64 path (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cov-mocks", type=int, default=512, help="mocks of the covariance run (0: skip)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false", help="keep the default launch shapes")
    ap.add_argument("--lanes", type=int, default=3, help="plans on streams of their own taking mocks in turn")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="also run one device copy of BYTES (scales HBM counters under rocprofv3 --pmc)")
    run(ap.parse_args())


if __name__ == "__main__":
    main()
