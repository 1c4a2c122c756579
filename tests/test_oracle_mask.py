"""Pins the oracle's mask handling (run_piv(...; mask, mask_threshold)).

* the masked window load (load_windows!, src/correlators.jl:214-227): a plane
  computed through process_windows with a pixel mask equals the direct linear
  correlation (scipy) of windows formed by definition — mean over the valid
  pixels, masked pixels zero, apodized — times the gain;
* the masked Wieneke statistics (load_uncertainty_windows!, src/uncertainty.jl:
  51-79) equal the unmasked statistics of windows whose masked pixels hold the
  valid mean (they contribute exactly zero either way);
* the grid mask (pass_grid, src/pipeline.jl:542-549) equals the window and
  search-area fractions counted with numpy;
* run_piv with a mask reproduces the reference's test scenes
  (test/test_masking.jl:69-125): masked cells NaN and never outliers, the
  unmasked field unaffected by a garbage region, mask_threshold = 1 masking
  fewer cells, multi-pass predictor filling keeping the field finite.
"""
import numpy as np
import pytest
from scipy.signal import correlate2d

from conftest import particle_pair


def P(hh, **kw):
    return hh.PIVParameters(**kw).to_struct()


def direct_masked_plane(A, B, m, apod, gain, M):
    valid = ~m
    a = np.where(valid, apod * (A - A[valid].mean()), 0.0)
    b = np.where(valid, apod * (B - B[valid].mean()), 0.0)
    full = correlate2d(b, a, mode="full")
    w = A.shape[0]
    R = np.zeros((M, M))
    H = M // 2
    for l1 in range(-(w - 1), w):
        for l2 in range(-(w - 1), w):
            R[l1 + H, l2 + H] = abs(full[l1 + w - 1, l2 + w - 1])
    return R * gain


@pytest.mark.parametrize("w,frac", [(16, 0.2), (32, 0.45)])
def test_masked_plane_equals_direct_correlation(orc, hh, w, frac):
    rng = np.random.default_rng(w)
    A = rng.random((w, w))
    B = rng.random((w, w))
    m = rng.random((w, w)) < frac
    p = P(hh, window_size=w, overlap=w // 2, padding=True, apodization="gauss")
    jobs = np.array([[1], [1], [1], [1]], dtype=np.int32, order="F")
    r = orc.process_windows(p, A, B, jobs, 1, 1, planes=True, mask=m)
    R = r["planes"][:, :, 0]
    ap = orc.apodization(w, w, True)
    Rd = direct_masked_plane(A, B, m, ap, orc.gain(p), 2 * w)
    assert np.allclose(R, Rd, atol=1e-12 * Rd.max())
    # a clean window takes the unmasked path
    r0 = orc.process_windows(p, A, B, jobs, 1, 1, planes=True, mask=np.zeros_like(m))
    r1 = orc.process_windows(p, A, B, jobs, 1, 1, planes=True)
    assert np.array_equal(r0["planes"], r1["planes"])


def test_masked_uq_equals_mean_filled_windows(orc, hh):
    rng = np.random.default_rng(5)
    w = 16
    A, B = particle_pair((w, w), [(5.0, 6.0), (11.0, 4.0), (8.0, 12.0), (13.0, 13.0)], 0.6, -0.4)
    A = A + 0.05 * rng.random((w, w))
    B = B + 0.05 * rng.random((w, w))
    m = np.zeros((w, w), bool)
    m[:5, 9:] = True
    p = P(hh, window_size=w, overlap=8, padding=True, apodization="gauss", uncertainty=True)
    jobs = np.array([[1], [1], [1], [1]], dtype=np.int32, order="F")
    uu, uv, st = orc.uncertainty_sweep(p, A, B, jobs, 1, 1, stats=True, mask=m)
    Af, Bf = A.copy(), B.copy()
    Af[m] = A[~m].mean()
    Bf[m] = B[~m].mean()
    uu2, uv2, st2 = orc.uncertainty_sweep(p, Af, Bf, jobs, 1, 1, stats=True)
    assert np.allclose(st, st2, rtol=1e-9, atol=1e-12 * np.abs(st2).max())
    assert uu[0, 0] == pytest.approx(uu2[0, 0], rel=1e-9) and uv[0, 0] == pytest.approx(uv2[0, 0], rel=1e-9)


@pytest.mark.parametrize("thr", [0.25, 0.5, 1.0])
def test_grid_mask_fractions(orc, hh, thr):
    rng = np.random.default_rng(11)
    nr, nc = 96, 80
    m = np.zeros((nr, nc), bool)
    m[20:60, 10:45] = True
    m |= rng.random((nr, nc)) < 0.05
    p = P(hh, window_size=16, overlap=8)
    gm = orc.grid_mask(p, nr, nc, m, thr)
    x, y, jobs = orc.pass_grid(p, nr, nc)
    want = np.zeros_like(gm)
    for k in range(jobs.shape[1]):
        gi, gj, rs, cs = jobs[:, k]
        want[gi - 1, gj - 1] = m[rs - 1:rs + 15, cs - 1:cs + 15].mean() >= thr
    assert np.array_equal(gm, want)


def _garbage_scene(n=128, seed=17):
    # test/test_masking.jl:69-82: 250 particles, shift (3, 2), a central
    # 32 x 32 region of uncorrelated garbage in both frames, masked
    rng = np.random.default_rng(seed)
    dv, du = 3.0, 2.0
    pos = [(rng.random() * (n + 20) - 10, rng.random() * (n + 20) - 10) for _ in range(250)]
    A, B = particle_pair((n, n), pos, dv, du)
    A[48:80, 48:80] = 3 * rng.random((32, 32))
    B[48:80, 48:80] = 3 * rng.random((32, 32))
    mask = np.zeros((n, n), bool)
    mask[48:80, 48:80] = True
    return A, B, mask, du, dv


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_run_piv_with_mask_reference_scene(orc, hh, prec):
    A, B, mask, du, dv = _garbage_scene()
    p = P(hh, window_size=32, overlap=16)
    r = orc.run_piv(A, B, [p], mask=mask, prec=prec)
    gm = r["mask"]
    assert gm.any() and not gm.all()
    assert np.all(np.isnan(r["u"][gm])) and np.all(np.isnan(r["peak_ratio"][gm]))
    assert not np.any(r["outliers"] & gm)
    good = r["u"][~gm]
    assert abs(np.median(good) - du) < 0.25 and abs(np.median(r["v"][~gm]) - dv) < 0.25
    assert np.count_nonzero(np.abs(good - du) < 0.5) >= 0.85 * good.size
    r1 = orc.run_piv(A, B, [p], mask=mask, mask_threshold=1.0, prec=prec)
    assert r1["mask"].sum() < gm.sum()
    r0 = orc.run_piv(A, B, [p], prec=prec)
    assert not r0["mask"].any()


def test_run_piv_with_mask_multipass(orc, hh):
    # test/test_masking.jl:103-109: predictor filling keeps the deformation
    # finite; masked cells stay NaN in the final result
    A, B, mask, du, dv = _garbage_scene()
    passes = [P(hh, window_size=64, overlap=32), P(hh, window_size=32, overlap=16)]
    r = orc.run_piv(A, B, passes, mask=mask)
    gm = r["mask"]
    assert gm.any()
    assert np.all(np.isnan(r["u"][gm])) and np.all(np.isfinite(r["u"][~gm]))
    assert abs(np.median(r["u"][~gm]) - du) < 0.25


def test_masked_single_window_recovers_shift(orc, hh):
    # test/test_masking.jl:57-67: a corrupted corner, masked, does not bias
    # the displacement of the whole-image window
    rng = np.random.default_rng(13)
    dv, du = 2.2, -1.7
    A, B = particle_pair((64, 64), [(40.0, 40.0), (20.0, 44.0)], dv, du)
    A[:24, :24] = 3 * rng.random((24, 24))
    B[:24, :24] = 3 * rng.random((24, 24))
    m = np.zeros((64, 64), bool)
    m[:24, :24] = True
    p = P(hh, window_size=64, overlap=0)
    jobs = np.array([[1], [1], [1], [1]], dtype=np.int32, order="F")
    r = orc.process_windows(p, A, B, jobs, 1, 1, mask=m)
    assert r["u"][0, 0] == pytest.approx(du, abs=0.1) and r["v"][0, 0] == pytest.approx(dv, abs=0.1)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_ensemble_one_pair_with_mask_is_run_piv(orc, hh, prec):
    # one pair: each ensemble pass is the single-pair pass, masks included
    # (src/ensemble.jl:154,254-264 against src/pipeline.jl:655-832)
    A, B, mask, du, dv = _garbage_scene()
    passes = [p.to_struct() for p in hh.multipass_parameters([32, 16], padding=True, apodization="gauss",
                                                             final=dict(uncertainty=True))]
    e = orc.run_piv_ensemble([(A, B)], passes, prec=prec, mask=mask)
    r = orc.run_piv(A, B, passes, prec=prec, mask=mask)
    assert e["mask"].any()
    for k in ("u", "v", "peak_ratio", "correlation_moment", "uncertainty_u", "uncertainty_v", "outliers", "mask"):
        assert np.array_equal(e[k], r[k], equal_nan=True), k


def test_ensemble_masked_plane_sums(orc, hh):
    # the masked ensemble sums each unmasked window's masked-load planes
    # over the pairs; masked cells are NaN and never outliers
    from conftest import low_snr_pairs

    pairs = low_snr_pairs(n=64, npairs=3)
    m = np.zeros((64, 64), bool)
    m[10:40, 20:50] = True
    p = P(hh, window_size=32, overlap=16)
    e = orc.run_piv_ensemble(pairs, [p], prec="f32", planes=True, mask=m)
    gm = orc.grid_mask(p, 64, 64, m)
    assert np.array_equal(e["mask"], gm) and gm.any() and not gm.all()
    x, y, jobs = orc.pass_grid(p, 64, 64)
    keep = [k for k in range(jobs.shape[1]) if not gm[jobs[0, k] - 1, jobs[1, k] - 1]]
    jobs = np.asfortranarray(jobs[:, keep])
    total = 0
    for A, B in pairs:
        r = orc.process_windows(p, A.astype(np.float32), B.astype(np.float32), jobs, y.size, x.size, prec="f32",
                                planes=True, mask=m)
        total = total + r["planes"].astype(np.float64)
    assert np.array_equal(e["planes"], total)
    assert np.all(np.isnan(e["u"][gm])) and not e["outliers"][gm].any()
    assert np.all(np.isfinite(e["u"][~gm]))
