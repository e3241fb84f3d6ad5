#!/usr/bin/env python3
"""Learned-reconstruction run: train the correction network on generated
mocks and compare it with the plain reconstruction on held-out mocks.

Workload (defaults): Zel'dovich tracers from a 512^3 lattice, L = 2000 Mpc/h,
4e6 tracers per mock selected with Lagrangian bias b_L = b - 1 = 1 (Eulerian
b = 2), redshift-space factor f = 0.76 along z, P(k) = 2e4 (k/0.05)^-1.5
exp(-(k/0.5)^2) [1 + 0.05 sin(2 pi k / 0.0628)]; the plain reconstruction
(b1 = 2, Gaussian smoothing R = 15 Mpc/h) and the network on a 256^3 mesh;
256 training mocks (kept resident in HBM), 64 held-out mocks.

Reported per held-out mock and averaged:
  - displacement error: rms over tracers of |S - T|, S the predicted shift,
    T the true one (the shift back to the tracer's lattice site);
  - post-reconstruction P_0: the tracers moved by S (plain, learned, true)
    and measured on a 256^3 box mesh; the BAO amplitude A of
    P_0 / smooth(P_0) - 1 projected on the input wiggle 0.05 sin(2 pi k / 0.0628)
    over 0.03 < k < 0.3 h/Mpc (A = 1: undamped);
  - cost: ms per mock of the plain reconstruction and of the learned one.

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

PK = dict(amp=2.0e4, k0=0.05, slope=-1.5, kcut=0.5, bao_amp=0.05, bao_period=0.0628)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bao_amplitude(k, p, kmin=0.03, kmax=0.3):
    """A from P / smooth - 1 = A 0.05 sin(2 pi k / 0.0628), smooth a quartic in log k of log P."""
    s = (k > kmin) & (k < kmax) & (p > 0)
    x, y = np.log(k[s]), np.log(p[s])
    w = 0.05 * np.sin(2.0 * np.pi * k[s] / PK["bao_period"])
    # joint fit of the smooth part and the wiggle (linear least squares in log space)
    X = np.stack([x ** j for j in range(5)] + [w], 1)
    coef = np.linalg.lstsq(X, y, rcond=None)[0]
    return float(coef[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--side", type=int, default=512, help="lattice per axis")
    ap.add_argument("--mesh", type=int, default=256, help="reconstruction / network mesh per axis")
    ap.add_argument("--boxsize", type=float, default=2000.0)
    ap.add_argument("--tracers", type=float, default=4e6)
    ap.add_argument("--train", type=int, default=256)
    ap.add_argument("--test", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--pk-mesh", type=int, default=256)
    ap.add_argument("--fine-mesh", type=int, default=512, help="mesh of the costlier plain reconstruction (0: skip)")
    ap.add_argument("--md", default=None)
    args = ap.parse_args()

    ctx = hm.Context(0)
    cfg = hm.mock_config(args.side, args.boxsize, f_growth=0.76, **PK)
    L = args.boxsize

    def gen(seed):
        return hm.mock_zeldovich_biased(ctx, cfg, seed, args.tracers, 1.0)

    t = time.perf_counter()
    train = [gen(10_000 + i) for i in range(args.train)]
    ctx.sync()
    gen_ms = (time.perf_counter() - t) * 1000.0 / max(args.train, 1)
    log(f"{args.train} training mocks generated, {gen_ms:.1f} ms each, "
        f"{np.mean([m[2] for m in train]):.0f} tracers on average")

    model = hm.LRecon(ctx, args.mesh, L, b1=2.0, RG=15.0, psi_scale=10.0, lr=args.lr, init_seed=1)
    hist = []
    t = time.perf_counter()
    for ep in range(args.epochs):
        order = np.random.default_rng(ep).permutation(args.train)
        l0s, l1s = [], []
        for i in order:
            l0, l1 = model.train_step(train[i][0], train[i][1])
            l0s.append(l0)
            l1s.append(l1)
        hist.append((float(np.sqrt(np.mean(l0s))), float(np.sqrt(np.mean(l1s)))))
        log(f"epoch {ep}: rms plain {hist[-1][0]:.3f}, learned (during training) {hist[-1][1]:.3f} Mpc/h")
    ctx.sync()
    train_ms = (time.perf_counter() - t) * 1000.0 / max(args.train * args.epochs, 1)
    for m in train:
        m[0].free()
        m[1].free()
    train = None

    plan = hm.BoxPlan(ctx, args.pk_mesh, L, "TSC", n_kbin=60, kmin=0.005, kmax=0.305, pk_ells=(0,), bk=False,
                      precision=hm.FP32)
    kc = 0.005 + 0.005 * (np.arange(60) + 0.5)
    rms = {"plain": [], "learned": []}
    fine = hm.LRecon(ctx, args.fine_mesh, L, b1=2.0, RG=15.0, psi_scale=10.0) if args.fine_mesh else None
    if fine:
        rms["plain_fine"] = []
    amp = {"pre": [], "plain": [], "learned": [], "true": []}
    cost = {"plain": 0.0, "learned": 0.0}
    for j in range(args.test):
        pos, shift, n = gen(90_000 + j)
        sp, sq = model.predict(pos)
        # cost: the plain reconstruction alone (features + gather) and the learned one (+ network)
        ctx.sync()
        t = time.perf_counter()
        model.predict(pos)
        ctx.sync()
        cost["learned"] += time.perf_counter() - t
        P, T = pos.download(), shift.download()
        S0, S1 = sp.download(), sq.download()
        if fine:
            fp, fq = fine.predict(pos)
            rms["plain_fine"].append(float(np.sqrt(np.mean(np.sum((fp.download() - T) ** 2, 1)))))
            fp.free()
            fq.free()
        rms["plain"].append(float(np.sqrt(np.mean(np.sum((S0 - T) ** 2, 1)))))
        rms["learned"].append(float(np.sqrt(np.mean(np.sum((S1 - T) ** 2, 1)))))
        for key, S in (("pre", None), ("plain", S0), ("learned", S1), ("true", T)):
            q = P if S is None else np.mod(P + S, L)
            q[q >= L] -= L
            r = plan.execute(q)
            amp[key].append(bao_amplitude(kc, r["pk"][0, :, 0]))
        for a in (pos, shift, sp, sq):
            a.free()
    # the plain reconstruction's cost, timed alone
    pos, shift, n = gen(99_999)
    feat = model.features(pos)
    ctx.sync()
    t = time.perf_counter()
    for _ in range(5):
        model.features(pos)
    ctx.sync()
    cost["plain"] = (time.perf_counter() - t) / 5
    if fine:
        fine.features(pos)
        ctx.sync()
        t = time.perf_counter()
        for _ in range(5):
            fine.features(pos)
        ctx.sync()
        cost["plain_fine"] = (time.perf_counter() - t) / 5
    out = {
        "workload": f"learned reconstruction, {args.side}^3 lattice, L = {L:g}, {args.tracers:.0f} tracers, b = 2, "
                    f"f = 0.76, R = 15; network on {args.mesh}^3; {args.train} training and {args.test} held-out mocks",
        "epochs": args.epochs,
        "train_history_rms": hist,
        "held_out_rms_plain": round(float(np.mean(rms["plain"])), 4),
        "held_out_rms_learned": round(float(np.mean(rms["learned"])), 4),
        "held_out_rms_gain": round(1.0 - float(np.mean(rms["learned"])) / float(np.mean(rms["plain"])), 4),
        "learned_better_on": f"{int(np.sum(np.array(rms['learned']) < np.array(rms['plain'])))} of {args.test}",
        "held_out_rms_plain_fine": round(float(np.mean(rms["plain_fine"])), 4) if fine else None,
        "fine_mesh": args.fine_mesh,
        "bao_amplitude": {k: round(float(np.mean(v)), 3) for k, v in amp.items()},
        "bao_amplitude_sem": {k: round(float(np.std(v) / np.sqrt(len(v))), 3) for k, v in amp.items()},
        "ms_per_mock": {"generate": round(gen_ms, 2), "train_step": round(train_ms, 2),
                        "plain_reconstruction": round(cost["plain"] * 1000.0 / 1.0, 2),
                        "learned_reconstruction": round(cost["learned"] * 1000.0 / max(args.test, 1), 2),
                        "plain_reconstruction_fine_mesh": round(cost.get("plain_fine", 0.0) * 1000.0, 2)},
    }
    print(json.dumps(out), flush=True)
    if args.md:
        with open(args.md, "w") as f:
            f.write("# Learned reconstruction run\n\n```\n" + json.dumps(out, indent=1) + "\n```\n")
    model.close()
    plan.close()
    ctx.close()


if __name__ == "__main__":
    main()
