"""Batched pipeline: each mock's data vector equals what the per-mock plan
returns (P_l real parts, then Re(B_000 - SN) on a <= b), bit for bit, for
mocks of uneven size, eagerly and replayed from the captured graph (with
eager assignment work between replays); the covariance equals numpy.cov of
the vectors (the N - 1 estimator, calc_data_and_cov.py:287)."""
import ctypes as ct

import numpy as np
import pytest

import hitomi_amd as hm

pytestmark = pytest.mark.gpu

N_SIDE, L = 32, 600.0
PLAN = dict(n_kbin=6, kmin=0.02, kmax=0.14, pk_ells=(0, 2, 4), bk=True)
SIZES = [32768, 21000, 26500, 17000, 30000, 24000]


def mocks(ctx):
    cfg = hm.mock_config(N_SIDE, L, amp=2.0e4, k0=0.05, slope=-1.5, kcut=0.5, bao_amp=0.05)
    return [hm.mock_zeldovich(ctx, cfg, 100 + i, n) for i, n in enumerate(SIZES)]


def per_mock_vector(res, nk, npk):
    v = [res["pk"][e, :, 0] for e in range(npk)]
    iu = np.triu_indices(nk)
    v.append(res["bk"][..., 0][iu])
    return np.concatenate(v)


@pytest.mark.parametrize("prec,graph,lanes", [(hm.FP32, True, 1), (hm.FP32, False, 1), (hm.FP64, True, 1),
                                              (hm.FP32, True, 3), (hm.FP32, False, 2)])
def test_batch_equals_per_mock_bitwise(ctx, prec, graph, lanes):
    plain = prec == hm.FP64
    kw = dict(PLAN, precision=prec, regrid=not plain, prune=not plain)
    pos = mocks(ctx)
    plan = hm.BoxPlan(ctx, 64, L, "TSC", **kw)
    ref = np.array([per_mock_vector(plan.execute(p), PLAN["n_kbin"], 3) for p in pos])
    plan.close()
    b = hm.Batch(ctx, 64, L, max_mocks=16, use_graph=graph, lanes=lanes, **kw)
    assert b.vec_len == 3 * 6 + 21 and b.cov_len == 18
    for p in pos:
        b.push(p)
    assert b.count == len(pos)
    got = b.vectors()
    assert np.array_equal(got, ref)
    # after a reset the replayed graph gives the same vectors again
    b.reset()
    for p in pos[:3]:
        b.push(p)
    assert np.array_equal(b.vectors(), ref[:3])
    b.close()


@pytest.mark.parametrize("with_bk,lanes", [(False, 1), (True, 1), (True, 3)])
def test_batch_covariance_matches_numpy(ctx, with_bk, lanes):
    pos = mocks(ctx)
    b = hm.Batch(ctx, 64, L, max_mocks=4, cov_with_bk=with_bk, lanes=lanes, **PLAN)
    for rep in range(2):
        for p in pos:
            b.push(p)
    for p in pos[:2]:
        b.push(p)
    assert b.count == 2 * len(pos) + 2
    stored = b.vectors()
    assert stored.shape == (4, b.vec_len)  # only max_mocks vectors are kept
    # the same vectors again from a fresh plan, for numpy
    plan = hm.BoxPlan(ctx, 64, L, "TSC", **PLAN)
    one = [per_mock_vector(plan.execute(p), 6, 3) for p in pos]
    plan.close()
    vecs = np.array(one * 2 + one[:2])
    assert np.array_equal(stored, vecs[:4])
    x = vecs[:, : b.cov_len]
    mean, cov = b.covariance()
    ref = np.cov(x.T)
    sd = np.sqrt(np.diag(ref))
    assert np.abs(mean - x.mean(0)).max() <= 1e-12 * np.abs(x).max()
    assert np.abs(cov - ref).max() <= 1e-10 * sd.max() ** 2
    assert np.allclose(cov, cov.T, rtol=0, atol=1e-12 * sd.max() ** 2)
    b.close()


def test_batch_needs_two_mocks_and_rejects_unsupported(ctx):
    b = hm.Batch(ctx, 64, L, max_mocks=2, **PLAN)
    b.push(mocks(ctx)[0])
    with pytest.raises(hm.HitomiError):
        b.covariance()
    b.close()
    with pytest.raises(hm.HitomiError):
        _interlaced(ctx)


def _interlaced(ctx):
    cfg = hm.BatchConfig()
    cfg.box = hm.box_config(64, L, **dict(PLAN, bk=False))
    cfg.box.interlace = 1
    cfg.max_mocks = 2
    h = ctx.lib.hm_batch_create(ctx.h, ct.byref(cfg))
    if not h:
        ctx.raise_error("hm_batch_create")
    ctx.lib.hm_batch_destroy(ctx.h, h)


def test_lanes_refill_one_buffer(ctx):
    """A caller that regenerates every mock into one buffer on its own stream
    gets the same vectors with three lanes as with one."""
    cfg = hm.mock_config(N_SIDE, L, amp=2.0e4, k0=0.05, slope=-1.5, kcut=0.5, bao_amp=0.05)
    buf = ctx.empty((max(SIZES), 3), np.float64)
    out = []
    for lanes in (1, 3):
        b = hm.Batch(ctx, 64, L, max_mocks=16, lanes=lanes, **PLAN)
        for i, n in enumerate(SIZES):
            hm.mock_zeldovich(ctx, cfg, 100 + i, n, out=buf)
            b.push(buf, n)
        out.append(b.vectors())
        b.close()
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("lanes", [1, 3])
def test_lane_timing_busy_time(ctx, lanes):
    """hm_timing_get_busy: the union of the timed launches' intervals, merged
    over the lanes' streams.  One lane launches one kernel at a time, so the
    busy time equals the summed durations; with three lanes it lies between
    the sum over three and the sum."""
    pos = mocks(ctx)
    b = hm.Batch(ctx, 64, L, max_mocks=16, use_graph=False, lanes=lanes, **PLAN)
    for p in pos[:lanes]:
        b.push(p)
    b.collect_timing()
    ctx.timing(True)
    ctx.timing_select("bin_pk")
    ctx.timing_reset()
    for p in pos:
        b.push(p)
    b.collect_timing()
    total, n = ctx.kernel_time("bin_pk")
    busy = ctx.kernel_busy("bin_pk")
    ctx.timing(False)
    ctx.timing_select(None)
    b.close()
    assert n > 0 and n % len(pos) == 0
    assert 0.0 < busy <= total * (1.0 + 1e-6) + 1e-6
    if lanes == 1:
        assert busy >= total * (1.0 - 1e-6) - 1e-6
    else:
        assert busy >= total / lanes * (1.0 - 1e-6) - 1e-6


def test_covariance_workload_full_size_bitwise(ctx):
    """The covariance workload at its stated size: mocks of uneven size
    (Poisson around a Gamma-scattered mean of 1.5e6, truncated to
    [0.8e6, 2.5e6]) generated on the device from a 256^3 lattice, L = 1500,
    256^3 mesh, 20 bins 0.015-0.205, P_0, P_2, P_4 and B_000; the graph-replayed
    batch with autotuned launch shapes and three lanes gives each mock's
    vector bit for bit as the per-mock plan does, and the covariance equals
    numpy.cov of those vectors within 1e-10 of the largest variance."""
    from hitomi_amd import mocks as hmm

    sizes = hmm.poisson_sizes(5, 1_500_000, 800_000, 2_500_000, seed=77, scatter=0.25)
    cfg = hm.mock_config(256, 1500.0, amp=2.0e4, k0=0.05, slope=-1.5, kcut=0.5, bao_amp=0.05, bao_period=0.0628)
    kw = dict(n_kbin=20, kmin=0.015, kmax=0.205, pk_ells=(0, 2, 4), bk=True)
    pos = [hm.mock_zeldovich(ctx, cfg, 900 + i, int(n)) for i, n in enumerate(sizes)]
    plan = hm.BoxPlan(ctx, 256, 1500.0, "TSC", precision=hm.FP32, **kw)
    ref = np.array([per_mock_vector(plan.execute(p), 20, 3) for p in pos])
    plan.close()
    b = hm.Batch(ctx, 256, 1500.0, max_mocks=8, use_graph=True, lanes=3, precision=hm.FP32, **kw)
    b.autotune(pos[0], reps=1)
    for p in pos:
        b.push(p)
    got = b.vectors()
    assert np.array_equal(got, ref)
    mean, cov = b.covariance()
    x = ref[:, : b.cov_len]
    want = np.cov(x.T)
    assert cov.shape == (60, 60)
    assert np.abs(cov - want).max() <= 1e-10 * np.diag(want).max()
    assert np.abs(mean - x.mean(0)).max() <= 1e-12 * np.abs(x).max()
    b.close()
    for p in pos:
        p.free()
