"""Pins the oracle's enlarged search areas (search_area_size > window_size).

* the plane (load_windows! general path, src/correlators.jl:230-262, and
  restrict_search_lags!, :347-355) equals, on the lags the margins allow,
  the direct linear correlation of the apodized frame-A window (centred in
  the search frame) with the mean-subtracted, unapodized search area, and is
  zero elsewhere; no overlap gain applies (src/correlators.jl:73);
* run_piv reproduces the reference's search-area test
  (test/runtests.jl:66-103): the 11 x 11 grid at 24.5:8:104.5, integer
  medians (10, 3), the equal-area legacy path, and a search-only masked
  fringe dropping a node.
"""
import numpy as np
import pytest


def P(hh, **kw):
    return hh.PIVParameters(**kw).to_struct()


def direct_search_plane(A, B, apod, M):
    w, s = A.shape[0], B.shape[0]
    m, H = (s - w) // 2, M // 2
    a = apod * (A - A.mean())
    b = B - B.mean()
    R = np.zeros((M, M))
    for l1 in range(-m, m + 1):
        for l2 in range(-m, m + 1):
            R[H + l1, H + l2] = abs(np.sum(a * b[m + l1:m + l1 + w, m + l2:m + l2 + w]))
    return R


@pytest.mark.parametrize("w,s,padding,apod", [(16, 32, False, "gauss"), (16, 32, True, "gauss"),
                                              (16, 48, False, "none"), (8, 24, True, "none")])
def test_search_plane_equals_direct_correlation(orc, hh, w, s, padding, apod):
    rng = np.random.default_rng(w + s)
    A = rng.random((w, w))
    B = rng.random((s, s))
    p = P(hh, window_size=w, search_area_size=s, overlap=w // 2, padding=padding, apodization=apod)
    R = orc.correlation_plane(p, A, B)
    M = 2 * s if padding else s
    assert R.shape == (M, M)
    Rd = direct_search_plane(A, B, orc.apodization(w, w, apod == "gauss"), M)
    assert np.allclose(R, Rd, atol=1e-11 * Rd.max())
    # no overlap gain for an enlarged area
    assert not orc.gain(p).any()


def _reference_scene(seed=71):
    # test/runtests.jl:66-69: B is A moved by (3, 10) px
    rng = np.random.default_rng(seed)
    A = rng.random((128, 128))
    B = np.zeros((128, 128))
    B[3:, 10:] = A[:-3, :-10]
    return np.asfortranarray(A), np.asfortranarray(B)


def test_run_piv_search_area_reference_scene(orc, hh):
    A, B = _reference_scene()
    p = P(hh, window_size=16, search_area_size=48, overlap=8, subpixel_method="none", n_peaks=1, uod_enable=False)
    r = orc.run_piv(A, B, [p])
    assert r["u"].shape == (11, 11)
    assert np.array_equal(r["x"], np.arange(24.5, 105, 8)) and np.array_equal(r["y"], np.arange(24.5, 105, 8))
    assert np.median(r["u"]) == 10 and np.median(r["v"]) == 3
    # explicit equality is the legacy path
    p0 = P(hh, window_size=16, overlap=8, uod_enable=False)
    p1 = P(hh, window_size=16, search_area_size=16, overlap=8, uod_enable=False)
    r0, r1 = orc.run_piv(A, B, [p0]), orc.run_piv(A, B, [p1])
    assert np.array_equal(r0["u"], r1["u"], equal_nan=True) and np.array_equal(r0["v"], r1["v"], equal_nan=True)
    # a search-only masked fringe drops the node whose window is clean
    mask = np.zeros((128, 128), bool)
    mask[:16, :48] = True
    rm = orc.run_piv(A, B, [p], mask=mask, mask_threshold=0.3)
    assert rm["mask"][0, 0]


def test_search_area_uncertainty_uses_window_sized_b(orc, hh):
    # the uncertainty of an enlarged-search pass reads the window-sized B at
    # the window origin (src/pipeline.jl:989): the same as the equal-area pass
    A, B = _reference_scene(5)
    pe = P(hh, window_size=16, search_area_size=32, overlap=8, uncertainty=True)
    p0 = P(hh, window_size=16, overlap=8, uncertainty=True)
    x, y, jobs = orc.pass_grid(pe, 128, 128)
    ue = orc.uncertainty_sweep(pe, A, B, jobs, y.size, x.size)
    u0 = orc.uncertainty_sweep(p0, A, B, jobs, y.size, x.size)
    assert np.array_equal(ue[0], u0[0], equal_nan=True) and np.array_equal(ue[1], u0[1], equal_nan=True)
