"""Autotuner: the chosen launch shapes (wide column tiles of the FP32 passes
of length 256, 270 or 288, and 512; the P_l binning block size) leave every result
bit for bit unchanged, for the plan and for the graph-replayed batch."""
import numpy as np
import pytest

import hitomi_amd as hm

pytestmark = pytest.mark.gpu

KW = dict(n_kbin=20, kmin=0.015, kmax=0.205, pk_ells=(0, 2, 4), bk=True)


@pytest.fixture(scope="module")
def mock(ctx):
    cfg = hm.mock_config(128, 2000.0, bao_amp=0.05)
    return hm.mock_zeldovich(ctx, cfg, 77, 1_000_000)


def test_plan_shapes_do_not_change_results(ctx, mock):
    plan = hm.BoxPlan(ctx, 512, 2000.0, "TSC", precision=hm.FP32, **KW)
    assert plan.triple_mesh == (270, 270, 270)
    assert plan.shapes == (0, 256)
    ref = plan.execute(mock)
    for shapes in [(0b110, 64), (0b010, 128), (0b100, 256), (0b1000, 256), (0b1110, 128)]:
        plan.shapes = shapes
        assert plan.shapes == shapes
        out = plan.execute(mock)
        for k in ref:
            assert np.array_equal(out[k], ref[k]), (shapes, k)
    report = plan.autotune(mock, reps=1)
    print(report)
    assert "chosen" in report and report.count("ms;") == 24  # (270/288, 512, shell x pass) x 3 block sizes
    out = plan.execute(mock)
    for k in ref:
        assert np.array_equal(out[k], ref[k]), k
    plan.close()


def test_batch_autotune_keeps_vectors(ctx, mock):
    b = hm.Batch(ctx, 512, 2000.0, max_mocks=8, use_graph=True, **KW)
    for _ in range(3):
        b.push(mock)
    ref = b.vectors()
    assert np.array_equal(ref[0], ref[2])
    report = b.autotune(mock, reps=1)
    assert "chosen" in report
    b.reset()
    for _ in range(3):
        b.push(mock)
    assert np.array_equal(b.vectors(), ref)
    b.close()
