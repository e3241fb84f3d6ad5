"""Per-seam entry points (hm_grid_alloc ... hm_survey_volume) against the
oracle's restatement of the same reference member functions, on the
reference's complex FP64 grid layout, through the C-ABI.

Mode and point counts must be bit-exact; FP64 values within 1e-11 of the
largest magnitude (the FFTs and sums differ in rounding order only).
Meshes: anisotropic 12x16x20 (TSC), odd 15x9x14 (NGP), 16^3 (CIC) and
96x96x64 (TSC, several rows per lane in the binned sums)."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-11
MESHES = [((12, 16, 20), (300.0, 400.0, 500.0), "TSC"), ((15, 9, 14), (150.0, 90.0, 140.0), "NGP"),
          ((16, 16, 16), (500.0, 500.0, 500.0), "CIC")]
ASSIGN = {"NGP": po.NGP, "CIC": po.CIC, "TSC": po.TSC}


@pytest.fixture(scope="module")
def sm():
    from hitomi_amd import seams

    return seams


def _close(a, b, tol=TOL):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max() / scale
    assert err < tol, f"relative error {err:.3e}"


def _grids(n, L, seed):
    rng = np.random.default_rng(seed)
    s = tuple(n)
    return (rng.normal(size=s) + 1j * rng.normal(size=s), rng.normal(size=s) + 1j * rng.normal(size=s))


def _field(ctx, sm, n, L, a, arr=None):
    f = sm.DensityField(ctx, n, L, a)
    if arr is not None:
        f.upload(arr)
    return f


@pytest.mark.parametrize("n,L,a", MESHES)
def test_field_transforms_window(ctx, sm, n, L, a):
    rng = np.random.default_rng(11)
    npart = 3000
    pos = rng.random((npart, 3)) * np.asarray(L)
    w = rng.random(npart) + 0.5 + 1j * (rng.random(npart) - 0.5)
    p = po.param(n, L, ASSIGN[a])
    ref = po.calc_field(p, pos, np.stack([w.real, w.imag], 1))
    f = _field(ctx, sm, n, L, a)
    f.calcField(pos, w)
    got = f.download()
    assert np.abs(got - ref).max() <= 2.0 ** -30 * np.abs(ref).max()
    # continue from the oracle's own field so that each seam is compared alone
    f.upload(ref)
    f.calcFourierTransform()
    refk = po.fourier_transform(p, ref)
    _close(f.download(), refk)
    f.upload(refk)
    f.calcAssignmentFunctionCorrection()
    refw = po.assignment_correction(p, refk)
    _close(f.download(), refw)
    f.upload(refw)
    f.calcInverseFourierTransform()
    _close(f.download(), po.inverse_fourier_transform(p, refw))
    f.finalizeDensityField()


@pytest.mark.parametrize("n,L,a", MESHES + [((96, 96, 64), (1200.0, 1200.0, 800.0), "TSC")])
def test_power_spectrum(ctx, sm, n, L, a):
    g1, g2 = _grids(n, L, 3)
    p = po.param(n, L, ASSIGN[a])
    kb = np.linspace(0.02, 0.3, 8) if max(n) < 64 else np.linspace(0.01, 0.2, 20)
    p.n_kbin = kb.size
    d1, d2 = _field(ctx, sm, n, L, a, g1), _field(ctx, sm, n, L, a, g2)
    st = sm.TwoPointStatistics(ctx)
    sn = 3.5 - 1.25j
    for ell, m in [(0, 0), (2, 0), (2, 1), (1, -1), (4, 3)]:
        ref, nref = po.calc_power_spectrum(p, g1, g2, kb, sn, ell, m)
        st.calcPowerSpectrum(d1, d2, kb, sn, ell, m)
        assert np.array_equal(st.n_mode_pk, nref.astype(np.int64)), (ell, m)
        _close(st.pk, ref)


@pytest.mark.parametrize("n,L,a", MESHES + [((96, 96, 64), (1200.0, 1200.0, 800.0), "TSC")])
def test_correlation_function(ctx, sm, n, L, a):
    g1, g2 = _grids(n, L, 4)
    p = po.param(n, L, ASSIGN[a])
    rb = np.linspace(20.0, 120.0, 11) if max(n) < 64 else np.linspace(30.0, 150.0, 25)
    p.n_rbin = rb.size
    d1, d2 = _field(ctx, sm, n, L, a, g1), _field(ctx, sm, n, L, a, g2)
    st = sm.TwoPointStatistics(ctx)
    for ell, m in [(0, 0), (2, 0), (3, -2)]:
        ref, nref = po.calc_correlation_function(p, g1, g2, rb, 2.0 + 0.5j, ell, m)
        st.calcCorrelationFunction(d1, d2, rb, 2.0 + 0.5j, ell, m)
        assert np.array_equal(st.n_mode_xi, nref.astype(np.int64)), (ell, m)
        _close(st.xi, ref)


@pytest.mark.parametrize("n,L,a", MESHES)
def test_correlation_function_3pcf(ctx, sm, n, L, a):
    g1, g2 = _grids(n, L, 5)
    p = po.param(n, L, ASSIGN[a])
    rb = np.array([15.0, 30.0, 40.0, 55.0, 80.0, 100.0])  # non-uniform edges
    d1, d2 = _field(ctx, sm, n, L, a, g1), _field(ctx, sm, n, L, a, g2)
    st = sm.TwoPointStatistics(ctx)
    for l1, m1, l2, m2 in [(0, 0, 0, 0), (1, 1, 1, -1), (2, 0, 2, 1)]:
        ref, nref = po.corr_for_3pcf(p, g1, g2, rb, 1.5, l1, m1, l2, m2)
        st.calcCorrelationFunctionForThreePointFunction(d1, d2, rb, 1.5, l1, m1, l2, m2)
        assert np.array_equal(st.n_mode_xi, nref.astype(np.int64))
        _close(st.xi, ref)


@pytest.mark.parametrize("n,L,a", MESHES)
def test_shotnoise_field_ijk(ctx, sm, n, L, a):
    g1, g2 = _grids(n, L, 6)
    p = po.param(n, L, ASSIGN[a])
    d1, d2 = _field(ctx, sm, n, L, a, g1), _field(ctx, sm, n, L, a, g2)
    xi = _field(ctx, sm, n, L, a)
    sm.TwoPointStatistics(ctx).calcShotNoiseForBispectrum_ijk(d1, d2, 4.0 - 2.0j, xi)
    _close(xi.download(), po.shotnoise_bispectrum_ijk(p, g1, g2, 4.0 - 2.0j))


@pytest.mark.parametrize("n,L,a", MESHES)
def test_shell_and_bessel_inverse_transforms(ctx, sm, n, L, a):
    g1, _ = _grids(n, L, 7)
    p = po.param(n, L, ASSIGN[a])
    d = _field(ctx, sm, n, L, a, g1)
    out = _field(ctx, sm, n, L, a)
    for ell, m, kmag in [(0, 0, 0.05), (2, 1, 0.08), (1, -1, 0.004)]:
        ref, nref = po.ifft_for_bispectrum_ylm(p, g1, kmag, 0.02, ell, m)
        nm = out.calcInverseFourierTransformForBispectrum(d, kmag, 0.02, ell, m)
        assert nm == nref
        if nref == 0:  # an empty shell: the reference divides 0 by n_mode = 0 (field.hpp:395-398)
            assert np.isnan(out.download()).all() and np.isnan(ref).all()
        else:
            _close(out.download(), ref)
    for ell_j, ell, m, r in [(0, 0, 0, 35.0), (2, 2, -1, 60.0), (1, 1, 1, 12.5)]:
        ref = po.ifft_for_3pcf(p, g1, r, ell_j, ell, m)
        out.calcInverseFourierTransformForThreePointFunction(d, r, ell, m, ell_j=ell_j)
        _close(out.download(), ref)


def test_triple_sums_many_pairs(ctx, sm):
    n, L, a = (12, 16, 20), (300.0, 400.0, 500.0), "TSC"
    p = po.param(n, L, po.TSC)
    rng = np.random.default_rng(8)
    arrs = [rng.normal(size=n) + 1j * rng.normal(size=n) for _ in range(6)]
    f3 = rng.normal(size=n) + 1j * rng.normal(size=n)
    fields = [_field(ctx, sm, n, L, a, x) for x in arrs]
    F3 = _field(ctx, sm, n, L, a, f3)
    pairs = [(i, j) for i in range(6) for j in range(i, 6)]  # 21 pairs: two chunks
    got = sm.triple_sums(ctx, fields, pairs, F3)
    ref = np.array([po.triple_sum(p, arrs[i], arrs[j], f3) for i, j in pairs])
    _close(got, ref)


@pytest.mark.parametrize("l1,m1,l2,m2", [(0, 0, 0, 0), (1, 0, 1, 1), (2, -1, 1, 0)])
def test_sn_bessel_sweep(ctx, sm, l1, m1, l2, m2):
    n, L, a = (12, 16, 20), (300.0, 400.0, 500.0), "TSC"
    p = po.param(n, L, po.TSC)
    g1, _ = _grids(n, L, 9)
    xi = _field(ctx, sm, n, L, a, g1)
    k1 = np.full(18, 0.05)
    k2 = np.linspace(0.02, 0.2, 18)  # 18 pairs: two chunks
    got = sm.sn_bessel_sweep(ctx, xi, l1, m1, l2, m2, k1, k2)
    ref = np.array([po.sn_bessel_sweep(p, g1, l1, m1, l2, m2, u, v) for u, v in zip(k1, k2)])
    _close(got, ref, 1e-10)


def test_particle_sums_and_survey_volume(ctx, sm):
    rng = np.random.default_rng(10)
    N = 20000
    los = rng.normal(size=(N, 3))
    w = rng.random(N) + 0.2
    for ell, m, pw in [(0, 0, 2), (2, 0, 2), (2, -2, 3), (4, 1, 3), (1, 1, 1)]:
        got = sm.particle_sums(ctx, los, w, ell, m, pw)
        ref = po.particle_sum(los, w, ell, m, pw)
        assert abs(got - ref) < 1e-12 * max(abs(ref), np.sum(w ** pw) * 1e-3)
    assert sm.particle_sums(ctx, None, w, 0, 0, 2) == pytest.approx(np.sum(w * w), rel=1e-13)
    for n, L, a in MESHES:
        pos = rng.random((5000, 3)) * np.asarray(L) * 0.8
        ref = po.survey_volume(po.param(n, L, ASSIGN[a]), pos, w[:5000])
        got = sm.survey_volume(ctx, n, L, pos, w[:5000], a)
        assert abs(got - ref) < 1e-12 * ref, (got, ref)


def test_seam_errors(ctx, sm):
    from hitomi_amd import HitomiError

    n, L = (12, 16, 20), (300.0, 400.0, 500.0)
    d = _field(ctx, sm, n, L, "TSC")
    st = sm.TwoPointStatistics(ctx)
    with pytest.raises(HitomiError, match="n_kbin >= 2"):
        st.calcPowerSpectrum(d, d, [0.1], 0.0, 0, 0)
    with pytest.raises(HitomiError, match="ell <= 16"):
        st.calcCorrelationFunction(d, d, [20.0, 30.0], 0.0, 17, 0)
    with pytest.raises(HitomiError, match="distinct"):
        d.calcInverseFourierTransformForBispectrum(d, 0.05, 0.01)
    d.finalizeDensityField()
    with pytest.raises(HitomiError, match="finalised"):
        d.calcFourierTransform()
