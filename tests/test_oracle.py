"""Pins the CPU oracle (oracle/) before it is trusted as the GPU checker.

Every oracle function is checked against something that does not restate its
own formula: the reference's own known-answer tests (test/*.jl, cited per
test), closed forms, independent library routines (scipy natural splines,
scipy/numpy direct correlation, numpy median/quantile) or brute force on
tiny inputs.
"""
import math

import numpy as np
import pytest
from scipy.interpolate import CubicSpline
from scipy.signal import correlate2d

from conftest import add_particle, low_snr_pairs, particle_pair

PRECS = ["f64", "f32"]


def P(hh, **kw):
    return hh.PIVParameters(**kw).to_struct()


# ---------------------------------------------------------------- peak finding
@pytest.mark.parametrize("prec", PRECS)
def test_find_peaks_reference_kats(orc, prec):
    # test/test_peaks.jl:8-48
    R = np.zeros((32, 32))
    add_particle(R, (16.0, 16.0), 3.0)
    R[7, 23] = 0.6
    R[25, 5] = 0.3
    pk = orc.find_peaks(R, 3, prec=prec)
    assert [l for _, l in pk] == [(16, 16), (8, 24), (26, 6)]
    assert pk[0][0] == pytest.approx(1.0) and pk[1][0] == pytest.approx(0.6) and pk[2][0] == pytest.approx(0.3)
    only2 = np.zeros((16, 16))
    only2[7, 7] = 1.0
    only2[7, 8] = 0.9
    assert len(orc.find_peaks(only2, 3, prec=prec)) == 1
    nearby = np.zeros((16, 16))
    nearby[7, 7], nearby[7, 8], nearby[7, 9] = 1.0, 0.1, 0.8
    assert len(orc.find_peaks(nearby, 2, peak_finder="exclusion", prec=prec)) == 1
    rmax = orc.find_peaks(nearby, 2, prec=prec)
    assert [l for _, l in rmax] == [(8, 8), (8, 10)]
    p = P(orc_hh(), n_peaks=2)
    ratio = orc.analyze_plane(nearby, p, prec=prec)[2]
    assert ratio == pytest.approx(1 / 0.8)


def orc_hh():
    import __graft_entry__ as ge

    return ge.load_package()


@pytest.mark.parametrize("prec", PRECS)
def test_find_peaks_crafted_planes(orc, prec):
    # crafted planes of test/test_ka.jl:51-70; expected lists follow from the
    # documented rules (earlier column-major point wins equal values, plateau
    # keeps the last point, secondaries must be positive, NaN is never a max)
    nearby = np.zeros((16, 16))
    nearby[7, 7], nearby[7, 8], nearby[7, 9] = 1.0, 0.1, 0.8
    nearby[4, 2] = 0.6
    nearby[11, 12] = 0.6
    assert [l for _, l in orc.find_peaks(nearby, 4, prec=prec)] == [(8, 8), (8, 10), (5, 3), (12, 13)]
    assert orc.find_peaks(np.ones((7, 9)), 4, prec=prec) == [(1.0, (7, 9))]
    nonpos = np.full((8, 8), -2.0)
    nonpos[2, 3] = -1.0
    assert orc.find_peaks(nonpos, 4, prec=prec) == [(-1.0, (3, 4))]
    assert orc.find_peaks(np.full((8, 8), np.nan), 4, prec=prec) == []


def brute_regionalmax(R, k):
    """All strict/weak local maxima by the documented rule, sorted by
    (-value, column-major index), truncated: a sort-based restatement."""
    nr, nc = R.shape
    cands = []
    for j in range(nc):
        for i in range(nr):
            v = R[i, j]
            if np.isnan(v):
                continue
            ok = True
            for jj in range(max(0, j - 1), min(nc, j + 2)):
                for ii in range(max(0, i - 1), min(nr, i + 2)):
                    if ii == i and jj == j:
                        continue
                    w = R[ii, jj]
                    if ii + jj * nr < i + j * nr:
                        ok &= not (w > v)
                    else:
                        ok &= not (w >= v)
            if ok:
                cands.append((-v, i + j * nr, v, (i + 1, j + 1)))
    cands.sort()
    top = cands[:k]
    if not top:
        return []
    found = 1
    while found < len(top) and top[found][2] > 0:
        found += 1
    return [(c[2], c[3]) for c in top[:found]]


def test_find_peaks_brute_force(orc):
    rng = np.random.default_rng(5)
    for trial in range(300):
        nr, nc = rng.integers(3, 9, size=2)
        # coarse values create ties and plateaus
        R = np.round(rng.normal(size=(nr, nc)), 1)
        if trial % 7 == 0:
            R[rng.integers(nr), rng.integers(nc)] = np.nan
        k = int(rng.integers(1, 5))
        assert orc.find_peaks(R, k) == brute_regionalmax(R, k), (trial, R)


# ---------------------------------------------------------------- subpixel
def test_gauss3_exact_on_gaussian(orc):
    r0, c0 = 10.37, 11.81
    i = np.arange(1, 22)[:, None]
    j = np.arange(1, 22)[None, :]
    R = np.exp(-((i - r0) ** 2) / 3.1 - (j - c0) ** 2 / 1.7)
    pr, pc = np.unravel_index(np.argmax(R), R.shape)
    ref = orc.gauss3(R, pr + 1, pc + 1)
    assert abs(ref[0] - r0) < 1e-10 and abs(ref[1] - c0) < 1e-10


def test_gauss9_reference_kats(orc):
    # test/test_accuracy.jl:10-26
    r0, c0 = 17.3, 15.6
    i = np.arange(1, 34)[:, None]
    j = np.arange(1, 34)[None, :]
    R = np.exp(-(0.20 * (i - r0) ** 2 + 0.16 * (i - r0) * (j - c0) + 0.10 * (j - c0) ** 2))
    pr, pc = np.unravel_index(np.argmax(R), R.shape)
    ref9 = orc.gauss9(R, pr + 1, pc + 1)
    ref3 = orc.gauss3(R, pr + 1, pc + 1)
    err9 = math.hypot(ref9[0] - r0, ref9[1] - c0)
    err3 = math.hypot(ref3[0] - r0, ref3[1] - c0)
    assert err9 < 1e-8 and err9 < err3
    assert orc.gauss9(R, 1, 5) == orc.gauss3(R, 1, 5)
    flat = np.ones((9, 9))
    assert orc.gauss9(flat, 5, 5) == orc.gauss3(flat, 5, 5)


def test_moment_and_ratio_kats(orc, hh):
    # test/runtests.jl:544-564
    R = np.zeros((32, 32))
    R[15, 15] = 1.0
    R[7, 23] = 0.25
    res = orc.analyze_plane(R, P(hh))
    assert res[2] == 4.0
    narrow = np.zeros((33, 33))
    wide = np.zeros((33, 33))
    add_particle(narrow, (17.0, 17.0), 2.0)
    add_particle(wide, (17.0, 17.0), 6.0)
    mn, mw = orc.moment(narrow, 17.0, 17.0), orc.moment(wide, 17.0, 17.0)
    assert 0 < mn < mw
    assert math.isnan(orc.moment(narrow, float("nan"), float("nan")))
    # single nonzero sample: the moment is its distance to the peak position
    one = np.zeros((33, 33))
    one[16, 16] = 2.0
    assert orc.moment(one, 17.3, 16.8) == pytest.approx(math.sqrt(0.3 ** 2 + 0.2 ** 2), rel=1e-14)


# ---------------------------------------------------------------- correlation plane
def direct_plane(A, B, apod, gain, M):
    """Linear cross-correlation by definition (scipy), |.| * gain, with zero
    lag at M//2 (0-based) — no FFT involved."""
    a = apod * (A - A.mean())
    b = apod * (B - B.mean())
    # c[l] = sum_x a[x] b[x + l]
    full = correlate2d(b, a, mode="full")  # index (l + w - 1)
    w = A.shape[0]
    R = np.zeros((M, M))
    H = M // 2
    for l1 in range(-(w - 1), w):
        for l2 in range(-(w - 1), w):
            R[l1 + H, l2 + H] = abs(full[l1 + w - 1, l2 + w - 1])
    return R * gain


@pytest.mark.parametrize("w,apod", [(8, "gauss"), (16, "gauss"), (16, "none"), (32, "gauss")])
def test_plane_equals_direct_linear_correlation(orc, hh, w, apod):
    rng = np.random.default_rng(w)
    A = rng.random((w, w))
    B = rng.random((w, w))
    p = P(hh, window_size=w, overlap=w // 2, padding=True, apodization=apod)
    R = orc.correlation_plane(p, A, B)
    ap = orc.apodization(w, w, apod == "gauss")
    g = orc.gain(p)
    # gain = autocorrelation of the apodization by direct sums, thresholded
    ac = correlate2d(ap, ap, mode="full")
    acp = np.zeros((2 * w, 2 * w))
    acp[1:, 1:] = ac
    wmax = acp.max()
    g_direct = np.where(acp >= wmax / 2, wmax / np.where(acp > 0, acp, 1), 0.0)
    on = g_direct > 0
    assert np.array_equal(on, g > 0)
    assert np.allclose(g[on], g_direct[on], rtol=1e-12)
    Rd = direct_plane(A, B, ap, g, 2 * w)
    assert np.allclose(R, Rd, atol=1e-12 * Rd.max())


def test_sign_convention(orc, hh):
    # B is A shifted by (+dv, +du) -> peak at (+dv, +du) (src/correlators.jl:4-7)
    A, B = particle_pair((32, 32), [(16.0, 16.0), (10.0, 21.0), (22.0, 9.0)], 3.0, -2.0)
    p = P(hh, window_size=32, padding=True, apodization="gauss")
    R = orc.correlation_plane(p, A, B)
    res = orc.analyze_plane(R, p)
    assert res[0] == pytest.approx(-2.0, abs=0.1) and res[1] == pytest.approx(3.0, abs=0.1)


@pytest.mark.parametrize("kw", [dict(), dict(padding=True), dict(apodization="gauss"),
                                dict(padding=True, apodization="gauss"), dict(correlation_method="phase"),
                                dict(correlation_method="phase", padding=True)])
def test_known_displacement(orc, hh, kw):
    # test/runtests.jl:165-178,204-220
    for dv, du in ((2.2, 1.3), (-3.4, 2.6), (0.0, 0.0)):
        A, B = particle_pair((64, 64), [(32, 32)], dv, du)
        p = P(hh, window_size=64, **kw)
        res = orc.analyze_plane(orc.correlation_plane(p, A, B), p)
        assert res[0] == pytest.approx(du, abs=0.1) and res[1] == pytest.approx(dv, abs=0.1)


# ---------------------------------------------------------------- statistics helpers
def test_median_quantile_vs_numpy(orc):
    rng = np.random.default_rng(1)
    for n in range(1, 40):
        v = rng.normal(size=n)
        assert orc.median(v) == np.median(v)
        for p in (0.0, 0.3, 0.95, 1.0):
            assert orc.quantile(v, p) == pytest.approx(np.quantile(v, p, method="linear"), rel=1e-14, abs=1e-14)
    assert math.isnan(orc.median([1.0, float("nan"), 2.0]))


# ---------------------------------------------------------------- validation
def test_uod_reference_kats(orc):
    # test/runtests.jl:529-542 (scene regenerated with numpy's RNG)
    rng = np.random.default_rng(7)
    u = 3.0 + 0.05 * rng.standard_normal((8, 8))
    v = 0.05 * rng.standard_normal((8, 8))
    assert not orc.uod(u, v, 2.0, neighborhood=1).any()
    us = u.copy()
    us[3, 4] += 50.0
    f = orc.uod(us, v, 2.0, neighborhood=1)
    assert f[3, 4] and f.sum() <= 3


def brute_uod(u, v, thr, nb, eps):
    nr, nc = u.shape
    out = np.zeros(u.shape, bool)
    for c in range(nc):
        for r in range(nr):
            nu = [u[r2, c2] for c2 in range(max(0, c - nb), min(nc, c + nb + 1))
                  for r2 in range(max(0, r - nb), min(nr, r + nb + 1)) if (r2, c2) != (r, c)]
            nv = [v[r2, c2] for c2 in range(max(0, c - nb), min(nc, c + nb + 1))
                  for r2 in range(max(0, r - nb), min(nr, r + nb + 1)) if (r2, c2) != (r, c)]
            mu, mv = np.median(nu), np.median(nv)
            ru = abs(u[r, c] - mu) / (np.median(np.abs(np.array(nu) - mu)) + eps)
            rv = abs(v[r, c] - mv) / (np.median(np.abs(np.array(nv) - mv)) + eps)
            out[r, c] = ru > thr or rv > thr
    return out


def test_uod_brute_force(orc):
    rng = np.random.default_rng(3)
    for nb in (1, 2, 3):
        u = rng.normal(size=(11, 9))
        v = rng.normal(size=(11, 9))
        u[rng.random(u.shape) < 0.1] += 5
        assert np.array_equal(orc.uod(u, v, 2.0, nb, 0.1), brute_uod(u, v, 2.0, nb, 0.1))


def test_validation_pipeline_kats(orc, hh):
    # test/test_validation.jl:5-67
    u = np.ones((32, 32))
    v = np.zeros((32, 32))
    pr = np.full((32, 32), 2.0)
    cm = np.full((32, 32), 0.2)
    u[4, 4] = 10.0
    pr[9, 9] = 1.05
    cm[14, 14] = 1.0
    u[19, 19] = 5.0

    def run(validation):
        p = hh.PIVParameters(uod_enable=False, replace_outliers=False, n_peaks=1, validation=validation).to_struct()
        return orc.validate_and_replace(p, u, v, pr, cm)[2]

    f = run((("peak_ratio", 1.2),))
    assert f[9, 9] and f.sum() == 1
    f = run((("correlation_moment", 0.5),))
    assert f[14, 14] and f.sum() == 1
    f = run((("velocity_magnitude", {"min": 0.0, "max": 5.0}),))
    assert f[4, 4] and f.sum() == 1
    f = run((("uod", {"threshold": 2.0, "neighborhood_size": 1}),))
    assert f[19, 19] and f[4, 4] and f.sum() == 2
    f = run((("peak_ratio", 1.2), ("velocity_magnitude", {"min": 0.0, "max": 8.0}),
             ("uod", {"threshold": 2.0, "neighborhood_size": 1})))
    assert f[9, 9] and f[4, 4] and f[19, 19] and not f[14, 14] and f.sum() == 3


def test_replace_and_smooth_kats(orc):
    # test/runtests.jl:501-527
    u = np.tile(np.arange(1.0, 7.0), (6, 1))
    v = -u.copy()
    u[2, 3], v[2, 3] = 99.0, -99.0
    inv = np.zeros((6, 6), bool)
    inv[2, 3] = True
    ru, rv = orc.replace_vectors(u, v, inv)
    assert ru[2, 3] == 4.0 and rv[2, 3] == -4.0
    u2 = np.full((3, 3), 7.0)
    u2[1, 1] = 42.0
    inv2 = np.ones((3, 3), bool)
    inv2[0, 0] = False
    assert orc.replace_vectors(u2, np.zeros((3, 3)), inv2)[0][1, 1] == 42.0
    assert np.allclose(orc.smooth_field(np.full((5, 5), 3.0)), 3.0)
    lin = np.array([[i + 2 * j for j in range(1, 7)] for i in range(1, 7)], float)
    assert np.allclose(orc.smooth_field(lin)[1:5, 1:5], lin[1:5, 1:5])


def test_substitution(orc):
    u = np.full((7, 7), 3.0)
    v = np.zeros((7, 7))
    u[3, 3], v[3, 3] = -4.0, 5.0
    out = np.zeros((7, 7), bool)
    out[3, 3] = True
    au = np.full((7, 7, 2), np.nan)
    av = np.full((7, 7, 2), np.nan)
    au[3, 3] = [10.0, 3.05]
    av[3, 3] = [10.0, 0.02]
    nu, nv, no, n = orc.substitute(u, v, out, au, av)
    assert n == 1 and nu[3, 3] == 3.05 and nv[3, 3] == 0.02 and not no[3, 3]
    au[3, 3] = [10.0, 9.0]
    nu, nv, no, n = orc.substitute(u, v, out, au, av)
    assert n == 0 and no[3, 3] and nu[3, 3] == -4.0


# ---------------------------------------------------------------- deformation
def tensor_natural_spline(img, y, x):
    nr, nc = img.shape
    rows = np.array([CubicSpline(np.arange(1, nc + 1), img[i], bc_type="natural")(x) for i in range(nr)])
    return CubicSpline(np.arange(1, nr + 1), rows, bc_type="natural")(y)


def test_prefilter_equals_natural_spline(orc):
    rng = np.random.default_rng(11)
    img = rng.random((23, 17))
    coef = orc.prefilter(img)
    for yy, xx in rng.uniform(1, [23, 17], size=(200, 2)):
        assert orc.cubic(coef, 23, 17, yy, xx) == pytest.approx(tensor_natural_spline(img, yy, xx), abs=1e-12)
    for yy, xx in ((1.0, 1.0), (23.0, 17.0), (1.0, 17.0), (12.0, 1.0)):
        assert orc.cubic(coef, 23, 17, yy, xx) == pytest.approx(tensor_natural_spline(img, yy, xx), abs=1e-12)


def test_spline_invariants(orc):
    i = np.arange(1, 21)[:, None]
    j = np.arange(1, 16)[None, :]
    lin = 0.3 * i - 0.7 * j + 2.0
    c = orc.prefilter(lin)
    for yy, xx in ((1.0, 1.0), (1.5, 14.9), (20.0, 15.0), (7.25, 3.5)):
        assert orc.cubic(c, 20, 15, yy, xx) == pytest.approx(0.3 * yy - 0.7 * xx + 2.0, abs=1e-12)
    assert orc.cubic(c, 20, 15, 0.99, 3.0) == 0.0 and orc.cubic(c, 20, 15, 3.0, 15.01) == 0.0
    # a quadratic is not reproduced next to the border (natural, not not-a-knot)
    q = np.broadcast_to((i - 5.0) ** 2, (20, 15)).copy()
    cq = orc.prefilter(q)
    assert abs(orc.cubic(cq, 20, 15, 1.5, 3.0) - (1.5 - 5.0) ** 2) > 1e-3


def test_deform_identities(orc):
    rng = np.random.default_rng(2)
    img = rng.random((40, 36))
    cA = orc.prefilter(img)
    py = np.array([8.0, 20.0, 32.0])
    px = np.array([8.0, 18.0, 28.0])
    z = np.zeros((3, 3))
    wA, wB = orc.deform(cA, cA, 40, 36, py, px, z, z)
    assert np.allclose(wA, img, atol=1e-12) and np.allclose(wB, img, atol=1e-12)
    du, dv = np.full((3, 3), 4.0), np.full((3, 3), 2.0)
    wA, wB = orc.deform(cA, cA, 40, 36, py, px, du, dv)
    # warpA(r, c) = A(r - 1, c - 2), warpB(r, c) = B(r + 1, c + 2)
    assert np.allclose(wA[1:, 2:], img[:-1, :-2], atol=1e-12)
    assert np.allclose(wB[:-1, :-2], img[1:, 2:], atol=1e-12)


# ---------------------------------------------------------------- uncertainty
def brute_uq_sigma(wA, wB):
    """Wieneke eqs. (5)-(9) with the FULL +-4 covariance square, numpy."""
    out = []
    for A, B in ((wA, wB), (wA.T, wB.T)):
        nr, nc = A.shape
        dC = A[:, :-1] * B[:, 1:] - A[:, 1:] * B[:, :-1]
        C0 = (A[:, :-1] * B[:, :-1]).sum()
        Cp = (A[:, :-1] * B[:, 1:]).sum()
        Cm = (A[:, 1:] * B[:, :-1]).sum()
        pad = np.concatenate([dC[:, :1], dC, dC[:, -1:]], axis=1)
        dCs = 0.25 * (pad[:, :-2] + 2 * pad[:, 1:-1] + pad[:, 2:])
        d = dCs - dCs.mean()
        S = {}
        m = d.shape[1]
        for dr in range(-4, 5):
            for dc in range(-4, 5):
                s = 0.0
                for c in range(m):
                    for r in range(nr):
                        if 0 <= r + dr < nr and 0 <= c + dc < m:
                            s += d[r, c] * d[r + dr, c + dc]
                S[(dr, dc)] = s
        s2 = S[(0, 0)]
        for ring in range(1, 5):
            ringk = [k for k in S if max(abs(k[0]), abs(k[1])) == ring]
            if max(S[k] for k in ringk) < 0.05 * S[(0, 0)]:
                break
            s2 += sum(S[k] for k in ringk)  # full ring = 2 x half ring
        sd = math.sqrt(max(s2, 0))
        Cpm = (Cp + Cm) / 2
        lo, hi = Cpm - sd / 2, Cpm + sd / 2
        if not (C0 > 0 and lo > 0):
            out.append(float("nan"))
            continue
        den = 2 * math.log(C0) - math.log(hi) - math.log(lo)
        out.append((math.log(hi) - math.log(lo)) / (2 * den) if den > 0 else float("nan"))
    return out


def test_uq_brute_force_and_identity(orc, hh):
    rng = np.random.default_rng(9)
    A, B = particle_pair((40, 40), [tuple(rng.uniform(1, 40, 2)) for _ in range(60)], 0.3, 0.4, diameter=2.5)
    A = A + 0.02 * rng.standard_normal(A.shape)
    B = B + 0.02 * rng.standard_normal(B.shape)
    p = P(hh, window_size=16, overlap=8, padding=True, apodization="gauss", uncertainty=True)
    x, y, jobs = orc.pass_grid(p, 40, 40)
    uu, uv = orc.uncertainty_sweep(p, A, B, jobs, y.size, x.size)
    ap = orc.apodization(16, 16, True)
    for k in range(jobs.shape[1]):
        gi, gj, rs, cs = jobs[:, k]
        sa = A[rs - 1:rs + 15, cs - 1:cs + 15]
        sb = B[rs - 1:rs + 15, cs - 1:cs + 15]
        wa, wb = ap * (sa - sa.mean()), ap * (sb - sb.mean())
        su, sv = brute_uq_sigma(wa, wb)
        assert uu[gi - 1, gj - 1] == pytest.approx(su, rel=1e-9, nan_ok=True)
        assert uv[gi - 1, gj - 1] == pytest.approx(sv, rel=1e-9, nan_ok=True)
    # identical windows: sigma_dC = 0 -> sigma = 0
    uu0, uv0 = orc.uncertainty_sweep(p, A, A, jobs, y.size, x.size)
    assert np.all(uu0 == 0.0) and np.all(uv0 == 0.0)


# ---------------------------------------------------------------- pipeline
def test_run_piv_uniform_shift(orc, hh):
    # test/runtests.jl:229-257: padded + gauss: median within 0.05, all within 0.2
    rng = np.random.default_rng(42)
    pos = [tuple(rng.random(2) * 148 - 10) for _ in range(250)]
    A, B = particle_pair((128, 128), pos, 3.0, 2.0)
    p = P(hh, window_size=32, overlap=16, padding=True, apodization="gauss")
    for prec in PRECS:
        r = orc.run_piv(A, B, [p], prec=prec)
        assert abs(np.median(r["u"]) - 2.0) < 0.05 and abs(np.median(r["v"]) - 3.0) < 0.05
        assert np.all(np.abs(r["u"] - 2.0) < 0.2)
        assert np.all(r["peak_ratio"] > 1)


def shear_scene(seed):
    rng = np.random.default_rng(seed)
    n, a = 128, 1 / 16
    A = np.zeros((n, n))
    B = np.zeros((n, n))
    for _ in range(900):
        p = (rng.random() * (n + 20) - 10, rng.random() * (n + 20) - 10)
        add_particle(A, p, 3.0)
        add_particle(B, (p[0], p[1] + a * (p[0] - 64)), 3.0)
    return np.asfortranarray(A), np.asfortranarray(B), a


def test_run_piv_multipass_shear(orc, hh):
    # test/runtests.jl:338-361
    A, B, a = shear_scene(11)
    kw = dict(padding=True, apodization="gauss")
    single = orc.run_piv(A, B, [P(hh, window_size=16, overlap=8, **kw)])
    multi = orc.run_piv(A, B, [p.to_struct() for p in hh.multipass_parameters([64, 32, 16, 16], **kw)])
    truth = lambda r: a * (r["y"][:, None] - 64) + 0 * r["x"][None, :]
    rms = lambda r: math.sqrt(np.mean((r["u"] - truth(r)) ** 2))
    assert rms(multi) < 0.08 and rms(multi) < rms(single)
    assert math.sqrt(np.mean(multi["v"] ** 2)) < 0.05
    assert multi["outliers"].sum() <= 3


def test_iterated_equals_repeated(orc, hh):
    # test/runtests.jl:391-400: an iterated final stage with convergence_tol
    # = 0 is bitwise the schedule that repeats the final pass, UQ included
    A, B, _ = shear_scene(17)
    rep = orc.run_piv(A, B, [p.to_struct() for p in hh.multipass_parameters([32, 16, 16, 16], uncertainty=True)])
    it = orc.run_piv(A, B, [p.to_struct() for p in hh.multipass_parameters(
        [32, 16], uncertainty=True, final=dict(max_iterations=3, convergence_tol=0.0))])
    for k in ("u", "v", "peak_ratio", "correlation_moment", "uncertainty_u", "uncertainty_v", "outliers"):
        assert np.array_equal(rep[k], it[k], equal_nan=(k != "outliers")), k
    assert np.isfinite(it["uncertainty_u"]).any()
    e2 = orc.run_piv(A, B, [p.to_struct() for p in hh.multipass_parameters(
        [32, 16], final=dict(max_iterations=2, convergence_tol=0.0))])
    eb = orc.run_piv(A, B, [p.to_struct() for p in hh.multipass_parameters(
        [32, 16], final=dict(max_iterations=5, convergence_tol=1e6))])
    assert np.array_equal(e2["u"], eb["u"]) and np.array_equal(e2["v"], eb["v"])


def test_float32_matches_float64(orc, hh):
    # test/runtests.jl:295-336
    rng = np.random.default_rng(21)
    pos = [tuple(rng.random(2) * 148 - 10) for _ in range(250)]
    A, B = particle_pair((128, 128), pos, 3.0, 2.0)
    passes = [p.to_struct() for p in hh.multipass_parameters([64, 32])]
    r32 = orc.run_piv(A, B, passes, prec="f32")
    r64 = orc.run_piv(A, B, passes, prec="f64")
    assert np.max(np.abs(r32["u"] - r64["u"])) < 0.05


def test_uq_tracks_measured_error(orc, hh):
    # test/test_uncertainty.jl:30-52 (scene regenerated with numpy's RNG)
    rng = np.random.default_rng(1701)
    dv, du = 0.3, 0.6
    passes = [p.to_struct() for p in hh.multipass_parameters([32, 32], padding=True, apodization="gauss",
                                                             uncertainty=True)]
    med = []
    for noise in (0.02, 0.05, 0.1, 0.2):
        pos = [tuple(rng.random(2) * 180 - 10) for _ in range(420)]
        A, B = particle_pair((160, 160), pos, dv, du)
        A = A + noise * rng.standard_normal(A.shape)
        B = B + noise * rng.standard_normal(B.shape)
        r = orc.run_piv(A, B, passes)
        valid = ~r["outliers"]
        rms = math.sqrt((np.mean((r["u"][valid] - du) ** 2) + np.mean((r["v"][valid] - dv) ** 2)) / 2)
        s = r["uncertainty_u"][valid & np.isfinite(r["uncertainty_u"])]
        assert s.size >= 0.98 * valid.sum()
        assert np.all(s >= 0)
        assert 0.4 * rms < np.median(s) < 2 * rms
        med.append(np.median(s))
    assert med == sorted(med) and med[-1] > 4 * med[0]


def test_field_change_vs_numpy(orc):
    rng = np.random.default_rng(4)
    u, v = rng.normal(size=(9, 7)), rng.normal(size=(9, 7))
    pu, pv = u + 0.1 * rng.normal(size=u.shape), v + 0.1 * rng.normal(size=v.shape)
    ref = np.quantile(np.maximum(np.abs(u - pu), np.abs(v - pv)).ravel(), 0.95, method="linear")
    assert orc.field_change(u, v, pu, pv) == pytest.approx(ref, rel=1e-14)


# ---------------------------------------------------------------- ensemble
@pytest.mark.parametrize("prec", PRECS)
def test_ensemble_one_pair_is_run_piv(hh, orc, prec):
    # one pair: every ensemble pass is exactly the single-pair pass (the
    # plane sum of one plane is that plane), including the deforming passes,
    # the forced replacement of earlier passes and uncertainty on the last
    pair = hh.synthetic.particle_pair(128, 128, hh.synthetic.lamb_oseen_flow(64.5, 64.5, 20.0, 3.0), seed=7,
                                      dtype=np.float64)
    passes = [p.to_struct() for p in hh.multipass_parameters(
        [32, 16], padding=True, apodization="gauss", final=dict(uncertainty=True))]
    e = orc.run_piv_ensemble([(pair.A, pair.B)], passes, prec=prec)
    r = orc.run_piv(pair.A, pair.B, passes, prec=prec)
    for k in ("u", "v", "peak_ratio", "correlation_moment", "uncertainty_u", "uncertainty_v", "outliers"):
        assert np.array_equal(e[k], r[k], equal_nan=True), k


def test_ensemble_plane_sums(hh, orc):
    # accumulate_planes! adds each pair's correlation_plane! (src/ensemble.jl:317-343):
    # the sums equal the per-pair planes of process_windows! added in FP64
    pairs = low_snr_pairs(n=64, npairs=3)
    p = P(hh, window_size=32, overlap=16)
    e = orc.run_piv_ensemble(pairs, [p], prec="f32", planes=True)
    x, y, jobs = orc.pass_grid(p, 64, 64)
    total = 0
    for A, B in pairs:
        r = orc.process_windows(p, A.astype(np.float32), B.astype(np.float32), jobs, y.size, x.size, prec="f32",
                                planes=True)
        total = total + r["planes"].astype(np.float64)
    assert np.array_equal(e["planes"], total)
    # identical pairs: the sum of N copies is exactly N times the plane
    e4 = orc.run_piv_ensemble([pairs[0]] * 4, [p], prec="f32", planes=True)
    e1 = orc.run_piv_ensemble([pairs[0]], [p], prec="f32", planes=True)
    assert np.array_equal(e4["planes"], 4 * e1["planes"])
    assert np.array_equal(e4["outliers"], e1["outliers"])
    fin = np.isfinite(e1["u"])
    assert np.abs(e4["u"][fin] - e1["u"][fin]).max() < 1e-5


def test_ensemble_low_snr_reference_bands(hh, orc):
    # behavioural bands of test/test_ensemble.jl:20-33 and :40-51 (unmasked)
    du, dv = 1.5, -2.0
    pairs = low_snr_pairs()
    p = P(hh, window_size=32, overlap=16)
    ens = orc.run_piv_ensemble(pairs, [p])
    assert np.all(np.abs(ens["u"] - du) < 0.5) and np.all(np.abs(ens["v"] - dv) < 0.5)
    assert abs(np.median(ens["u"]) - du) < 0.25 and abs(np.median(ens["v"]) - dv) < 0.25
    assert ens["outliers"].sum() <= 2
    single = orc.run_piv(*pairs[0], [p])
    err = lambda r: np.median(np.abs(r["u"][np.isfinite(r["u"])] - du))
    assert err(ens) <= err(single)
    multi = orc.run_piv_ensemble(pairs, [q.to_struct() for q in hh.multipass_parameters([64, 32])])
    assert abs(np.median(multi["u"]) - du) < 0.15 and abs(np.median(multi["v"]) - dv) < 0.15
    assert np.all(np.abs(multi["u"] - du) < 0.5)


# ---------------------------------------------------------------- adjoint
@pytest.mark.parametrize("interp", ["cubic", "bilinear"])
def test_energy_at_rest_is_windowed_correlation(hh, orc, interp):
    # zero node field: both interpolants reproduce the pixels, so E is the
    # apodized zero-lag covariance of every window, computed here directly
    rng = np.random.default_rng(2)
    A, B = rng.random((40, 36)), rng.random((40, 36))
    p = P(hh, window_size=8, overlap=4, apodization="gauss")
    py, px = np.array([4.5, 20.5, 36.5]), np.array([4.5, 18.5, 32.5])
    z = np.zeros((3, 3))
    E = orc.grad_energy(p, A, B, py, px, z, z, gradient=False, interp=interp)
    x, y, jobs = orc.pass_grid(p, 40, 36)
    # apodization_window (src/correlators.jl:18-25): centre (w+1)/2, sigma w/4
    i = np.arange(1, 9)
    g1 = np.exp(-((i - 4.5) ** 2) / (2 * 2.0 ** 2))
    apod = np.outer(g1, g1)
    ref = 0.0
    for k in range(jobs.shape[1]):
        rs, cs = jobs[2, k] - 1, jobs[3, k] - 1
        a = A[rs:rs + 8, cs:cs + 8]
        b = B[rs:rs + 8, cs:cs + 8]
        ref -= np.sum(apod * (a - a.mean()) * (b - b.mean()))
    assert E == pytest.approx(ref, rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("interp", ["cubic", "bilinear"])
def test_energy_gradient_central_differences(hh, orc, interp):
    # the adjoint against central differences of the FP64 energy (h = 1e-3 px)
    pair = hh.synthetic.particle_pair(96, 80, hh.synthetic.sine_shear_flow(2.0, 64.0), seed=5, dtype=np.float64)
    p = P(hh, window_size=16, overlap=12, apodization="gauss")
    py, px = np.arange(8.5, 96, 16.0), np.arange(8.5, 80, 16.0)
    # samples are zero outside the image, so E jumps where a border pixel's
    # displacement changes sign; positive fields keep every border sample on
    # one side of the edge, which leaves E smooth in all node values
    rng = np.random.default_rng(1)
    pu = rng.uniform(0.2, 1, (py.size, px.size))
    pv = rng.uniform(0.2, 1, (py.size, px.size))
    E, gu, gv = orc.grad_energy(p, pair.A, pair.B, py, px, pu, pv, interp=interp)
    h = 1e-3
    scale = max(np.abs(gu).max(), np.abs(gv).max())
    for (i, j) in [(0, 0), (2, 3), (py.size - 1, px.size - 1), (3, 0)]:
        for field in ("u", "v"):
            d = np.zeros_like(pu)
            d[i, j] = h
            args = lambda s: (pu + s * d, pv) if field == "u" else (pu, pv + s * d)
            fd = (orc.grad_energy(p, pair.A, pair.B, py, px, *args(1), gradient=False, interp=interp) -
                  orc.grad_energy(p, pair.A, pair.B, py, px, *args(-1), gradient=False, interp=interp)) / (2 * h)
            g = (gu if field == "u" else gv)[i, j]
            # bilinear sampling is piecewise linear: a sample crossing a pixel
            # line inside +-h adds an O(h) kink term
            tol = 1e-6 if interp == "cubic" else 2e-3
            assert abs(fd - g) <= tol * scale, (field, i, j, fd, g)


# ---------------------------------------------------------------- phase filter and node values
@pytest.mark.parametrize("w,padding,apod", [(16, False, "none"), (32, True, "gauss"), (24, True, "none"),
                                            (16, True, "gauss")])
def test_phase_plane_equals_numpy(orc, hh, w, padding, apod):
    # PhaseCorrelator (src/correlators.jl:110-127) and spectrum!
    # (:273-279): W = exp(-(fi^2 + fj^2) / (2 sigma^2)) over fftfreq(M, M)
    # bins in FFT order, sigma = min(M) / 8, P = W s / (|s| + eps), s =
    # conj(FA) FB; |ifft| fftshifted to M/2 + 1, times the overlap gain when
    # padded. Restated here with numpy's FFT and the closed-form apodization.
    rng = np.random.default_rng(w + padding)
    A = rng.random((w, w))
    B = np.roll(A, (2, -1), axis=(0, 1)) + 0.3 * rng.random((w, w))
    p = P(hh, window_size=w, overlap=w // 2, padding=padding, apodization=apod, correlation_method="phase")
    R = orc.correlation_plane(p, A, B)
    M = 2 * w if padding else w
    i = np.arange(1, w + 1)
    c = (w + 1) / 2
    s = w / 4
    ap = np.exp(-((i[:, None] - c) ** 2 / (2 * s * s) + (i[None, :] - c) ** 2 / (2 * s * s))) if apod == "gauss" \
        else np.ones((w, w))
    a = np.zeros((M, M))
    b = np.zeros((M, M))
    a[:w, :w] = ap * (A - A.mean())
    b[:w, :w] = ap * (B - B.mean())
    sp = np.conj(np.fft.fft2(a)) * np.fft.fft2(b)
    f = np.fft.fftfreq(M, 1.0 / M)
    sigma = M / 8
    Wf = np.exp(-(f[:, None] ** 2 + f[None, :] ** 2) / (2 * sigma ** 2))
    Rn = np.fft.fftshift(np.abs(np.fft.ifft2(Wf * sp / (np.abs(sp) + np.finfo(np.float64).eps))))
    if padding:
        Rn = Rn * orc.gain(p)
    assert np.allclose(R, Rn, rtol=0, atol=1e-11 * Rn.max())
    # the filter's width and bin order matter: a sigma of min(M)/4, or the
    # filter in fftshifted order, changes the plane well above the tolerance
    for Wbad in (np.exp(-(f[:, None] ** 2 + f[None, :] ** 2) / (2 * (M / 4) ** 2)), np.fft.fftshift(Wf)):
        Rb = np.fft.fftshift(np.abs(np.fft.ifft2(Wbad * sp / (np.abs(sp) + np.finfo(np.float64).eps))))
        if padding:
            Rb = Rb * orc.gain(p)
        assert not np.allclose(R, Rb, rtol=0, atol=1e-6 * Rn.max())


@pytest.mark.parametrize("prec", PRECS)
def test_node_values_equal_linear_interpolation(orc, prec):
    # predictor_node_values (src/pipeline.jl:561-569): Gridded(Linear()) over
    # the (y, x) knots with Flat() extrapolation = scipy's linear
    # RegularGridInterpolator at coordinates clamped into the knot range, on
    # a non-linear node field and points inside, on and outside the knots
    from scipy.interpolate import RegularGridInterpolator

    rng = np.random.default_rng(5)
    py = 16.5 + 16.0 * np.arange(9)
    px = 24.5 + 12.0 * np.arange(11)
    Y, X = np.meshgrid(py, px, indexing="ij")
    pu = np.sin(X / 17.0) * np.cos(Y / 23.0) + 0.3 * rng.standard_normal(Y.shape)
    pv = (X - 60.0) ** 2 / 900.0 - Y / 50.0
    y = np.concatenate([[-5.0, 1.0, 16.5], rng.uniform(0, 170, 20), [144.5, 150.0, 400.0]])
    x = np.concatenate([[0.0, 24.5], rng.uniform(-10, 170, 25), [144.5, 160.0]])
    gu, gv = orc.node_values(py, px, pu, pv, y, x, prec=prec)
    yy = np.clip(y, py[0], py[-1])
    xx = np.clip(x, px[0], px[-1])
    pts = np.stack(np.meshgrid(yy, xx, indexing="ij"), -1).reshape(-1, 2)
    ru = RegularGridInterpolator((py, px), pu)(pts).reshape(y.size, x.size)
    rv = RegularGridInterpolator((py, px), pv)(pts).reshape(y.size, x.size)
    tol = 1e-12 if prec == "f64" else 2e-6
    assert np.allclose(gu, ru, rtol=0, atol=tol * 3) and np.allclose(gv, rv, rtol=0, atol=tol * 10)
    # linear extrapolation instead of Flat() would differ outside the knots
    lin = RegularGridInterpolator((py, px), pu, bounds_error=False, fill_value=None)
    pts_raw = np.stack(np.meshgrid(y, x, indexing="ij"), -1).reshape(-1, 2)
    assert not np.allclose(gu, lin(pts_raw).reshape(y.size, x.size), atol=1e-3)
