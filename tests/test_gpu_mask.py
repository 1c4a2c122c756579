"""Pixel masks on the device (run_piv(...; mask, mask_threshold),
src/pipeline.jl:361-449 with 527-555, 617-645, 655-832) against the oracle,
through the C ABI.

Bars: the grid mask bit-exact; masked cells NaN in u, v, peak ratio, moment
and both uncertainties, never outliers; outlier flags within the near-tie
allowance of the run_piv bars; u, v RMS < 5e-3 px over the cells valid in
both (1e-3 px max for a single pass); sigma median within 1 %; validation
with a mask bit-exact; a batch of masked pairs bitwise equal to run_piv on
each pair alone.
"""
import sys

import numpy as np
import pytest

from parity_bars import field_parity

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "peak_ratio", "correlation_moment")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def disk_scene(hh, n=256, seed=61, speckle=0.0):
    """A uniform-shift particle pair with a disk of uncorrelated garbage in
    both frames (the masked region), optionally with scattered masked pixels
    that leave most windows partly masked."""
    pair = hh.synthetic.particle_pair(n, n, hh.synthetic.uniform_flow(2.3, -1.7), seed, ppp=0.03, noise=0.02)
    rng = np.random.default_rng(seed + 1)
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    m = (i - 110.3) ** 2 + (j - 140.7) ** 2 <= 37.5 ** 2
    A, B = pair.A.copy(), pair.B.copy()
    A[m] = 3 * rng.random(m.sum())
    B[m] = 3 * rng.random(m.sum())
    if speckle:
        m = m | (rng.random((n, n)) < speckle)
    return np.asfortranarray(A, dtype=np.float32), np.asfortranarray(B, dtype=np.float32), m


def run_masked(hh, orc, ctx, A, B, m, passes, thr=0.5):
    ctx.set_mask(m, thr)
    try:
        ctx.set_pair(A, B)
        g = ctx.run_piv(passes)
    finally:
        ctx.set_mask(None)
    o = orc.run_piv(A.astype(np.float64), B.astype(np.float64), [p.to_struct() for p in passes], mask=m, uq_pert=True,
                    mask_threshold=thr)
    return g, o


def check_masked(g, o, rms_tol=5e-3, max_tol=None, uq=False):
    gm = g["mask"]
    assert np.array_equal(gm, o["mask"])
    assert gm.any() and not gm.all()
    keys = FIELDS + (("uncertainty_u", "uncertainty_v") if uq else ())
    for k in keys:
        assert np.all(np.isnan(g[k][gm])), k
    assert not np.any(g["outliers"] & gm)
    assert np.all(np.isfinite(g["u"][~gm]))
    st = field_parity(g, o, label="masked", rms_tol=rms_tol, tol_px=1e-3 if max_tol is None else max_tol,
                      uncertainty=uq)
    say("masked cells", int(gm.sum()), "of", gm.size)
    return st


@pytest.mark.parametrize("form", ["fft32", "mfma16"])
@pytest.mark.parametrize("speckle", [0.0, 0.01])
def test_run_piv_mask_multipass_vs_oracle(hh, orc, form, speckle):
    A, B, m = disk_scene(hh, speckle=speckle)
    passes = hh.multipass_parameters([32, 16], padding=True, apodization="gauss",
                                     final=dict(overlap=8, max_iterations=2, convergence_tol=0.0, uncertainty=True))
    say("mask multipass", form, "speckle", speckle)
    with hh.Context(0) as c:
        c.set_form(hh.HH_FORM_MFMA16 if form == "mfma16" else hh.HH_FORM_FFT32)
        g, o = run_masked(hh, orc, c, A, B, m, passes)
    check_masked(g, o, uq=True)
    good = ~g["mask"]
    assert abs(np.median(g["u"][good]) - 2.3) < 0.05 and abs(np.median(g["v"][good]) + 1.7) < 0.05


@pytest.mark.parametrize("thr", [0.5, 1.0])
def test_run_piv_mask_single_pass_vs_oracle(hh, orc, ctx, thr):
    # alternatives substituted around the mask (n_peaks = 3), replacement on
    A, B, m = disk_scene(hh, seed=67, speckle=0.004)
    p = hh.PIVParameters(window_size=32, overlap=16, padding=True, apodization="gauss", n_peaks=3,
                         replace_outliers=True)
    say("mask single pass thr", thr)
    g, o = run_masked(hh, orc, ctx, A, B, m, [p], thr=thr)
    assert np.array_equal(g["mask"], orc.grid_mask(p.to_struct(), *A.shape, m, thr))
    check_masked(g, o, max_tol=1e-3)


def test_validate_with_mask_bit_exact(hh, orc, ctx):
    rng = np.random.default_rng(23)
    ny, nx = 37, 41
    for trial, nb in enumerate((1, 2, 3)):
        u = (2.0 + 0.1 * rng.standard_normal((ny, nx))).astype(np.float32)
        v = (-1.0 + 0.1 * rng.standard_normal((ny, nx))).astype(np.float32)
        bad = rng.random((ny, nx)) < 0.08
        u[bad] += rng.uniform(-5, 5, bad.sum()).astype(np.float32)
        m = np.zeros((ny, nx), bool)
        m[10:19, 12:25] = True
        m |= rng.random((ny, nx)) < 0.03
        u[m] = np.nan
        v[m] = np.nan
        pr = rng.uniform(1.0, 3.0, (ny, nx)).astype(np.float32)
        pr[m] = np.nan
        cm = rng.uniform(0.1, 1.0, (ny, nx)).astype(np.float32)
        au = np.full((ny, nx, 2), np.nan, np.float32)
        av = np.full((ny, nx, 2), np.nan, np.float32)
        au[..., 0] = (2.0 + 0.05 * rng.standard_normal((ny, nx))).astype(np.float32)
        av[..., 0] = (-1.0 + 0.05 * rng.standard_normal((ny, nx))).astype(np.float32)
        p = hh.PIVParameters(uod_neighborhood=nb, min_peak_ratio=1.2, n_peaks=2)
        for force in (True, False):
            say("validate mask", trial, nb, force)
            gu, gv, gf = ctx.validate_and_replace(p, u, v, pr, cm, au, av, force_replace=force, mask=m)
            ou, ov, of = orc.validate_and_replace(p.to_struct(), u, v, pr, cm, au, av, force_replace=force,
                                                  mask=m.astype(np.uint8), prec="f32")
            assert not gf[m].any()
            assert np.array_equal(gf, of)
            assert np.array_equal(gu, ou, equal_nan=True) and np.array_equal(gv, ov, equal_nan=True)


def _same(a, b):
    keys = FIELDS + ("uncertainty_u", "uncertainty_v", "outliers", "mask")
    return all(np.array_equal(a[k], b[k], equal_nan=(k not in ("outliers", "mask"))) for k in keys)


@pytest.mark.parametrize("dynamic", [False, True])
def test_masked_batch_equals_single_runs(hh, dynamic):
    """A batch of equally sized pairs under one mask (run_piv_sequence with a
    mask) gives exactly what run_piv gives on each pair; with convergence_tol
    > 0 the pairs stop at different sweeps, so later sweeps correlate the
    masked windows of a subset of the pairs."""
    pairs = hh.synthetic.config_batch(npairs=4, seed0=205000, sizes=((192, 224),))
    m = np.zeros((192, 224), bool)
    m[60:130, 90:150] = True
    m[5:9, :] = True
    final = dict(overlap=8, max_iterations=4 if dynamic else 2, convergence_tol=0.02 if dynamic else 0.0,
                 uncertainty=True)
    passes = hh.multipass_parameters([32, 16], padding=True, apodization="gauss", final=final)
    say("masked batch, dynamic", dynamic)
    with hh.Context(0) as cb, hh.Context(0) as cs:
        cb.set_mask(m, 0.5)
        cs.set_mask(m, 0.5)
        cb.set_pairs([(p.A, p.B) for p in pairs])
        rb = cb.run_piv_batch(passes)
        for q, p in enumerate(pairs):
            cs.set_pair(p.A, p.B)
            rs = cs.run_piv(passes)
            assert rs["mask"].any()
            assert _same(rb[q], rs), f"pair {q}"
        if not dynamic:
            # the same mask again keeps the plan: the next run replays its graph
            n0 = cb.graph_replays()
            cb.set_mask(m.copy(), 0.5)
            rb2 = cb.run_piv_batch(passes)
            assert cb.graph_replays() == n0 + 1
            assert all(_same(a, b) for a, b in zip(rb, rb2))
        # removing the mask gives the unmasked result
        cb.set_mask(None)
        ru = cb.run_piv_batch(passes)
        assert not any(r["mask"].any() for r in ru)


def test_run_piv_sequence_with_mask(hh):
    pairs = hh.synthetic.config_batch(npairs=3, seed0=206000, sizes=((160, 192),))
    m = np.zeros((160, 192), bool)
    m[40:100, 50:120] = True
    p = hh.PIVParameters(window_size=32, overlap=16, padding=True, apodization="gauss")
    seq = hh.run_piv_sequence([(q.A, q.B) for q in pairs], p, mask=m, mask_threshold=0.4)
    for r, q in zip(seq, pairs):
        single = hh.run_piv(q.A, q.B, p, mask=m, mask_threshold=0.4)
        assert r.mask.any() and np.array_equal(r.mask, single.mask)
        assert np.array_equal(r.u, single.u, equal_nan=True)
        assert np.array_equal(r.outliers, single.outliers)


@pytest.mark.parametrize("form", ["fft32", "mfma16"])
def test_ensemble_mask_vs_oracle(hh, orc, form):
    """run_piv_ensemble with a mask (src/ensemble.jl:151-154,254-264): the
    low-SNR ensemble scene with a masked block and scattered masked pixels,
    two passes with pooled uncertainty on the last."""
    from conftest import low_snr_pairs

    pairs = low_snr_pairs(n=128, npairs=6)
    rng = np.random.default_rng(3)
    m = np.zeros((128, 128), bool)
    m[30:75, 60:110] = True
    m |= rng.random((128, 128)) < 0.01
    passes = hh.multipass_parameters([32, 16], padding=True, apodization="gauss", final=dict(uncertainty=True))
    say("ensemble mask", form)
    f32 = [(A.astype(np.float32), B.astype(np.float32)) for A, B in pairs]
    with hh.Context(0) as c:
        c.set_form(hh.HH_FORM_MFMA16 if form == "mfma16" else hh.HH_FORM_FFT32)
        c.set_mask(m, 0.5)
        g = c.run_piv_ensemble(f32, passes)
    o = orc.run_piv_ensemble(pairs, [p.to_struct() for p in passes], mask=m)
    check_masked(g, o, uq=True)
    if form == "fft32":
        # the public keyword interface gives the same result
        r = hh.run_piv_ensemble(f32, passes, mask=m)
        assert np.array_equal(r.mask, g["mask"]) and np.array_equal(r.u, g["u"], equal_nan=True)


def test_ensemble_one_pair_with_mask_matches_run_piv(hh, ctx):
    A, B, m = disk_scene(hh, n=192, seed=71, speckle=0.005)
    passes = hh.multipass_parameters([32, 16], padding=True, apodization="gauss", final=dict(uncertainty=True))
    ctx.set_mask(m, 0.5)
    try:
        ctx.set_pair(A, B)
        r = ctx.run_piv(passes)
        e = ctx.run_piv_ensemble([(A, B)], passes)
    finally:
        ctx.set_mask(None)
    assert r["mask"].any()
    for k in FIELDS + ("uncertainty_u", "uncertainty_v", "outliers", "mask"):
        fin = np.isfinite(r[k]) if r[k].dtype != bool else None
        if fin is None:
            assert np.array_equal(e[k], r[k]), k
        else:
            assert np.array_equal(np.isfinite(e[k]), fin), k
            assert np.abs(e[k][fin] - r[k][fin]).max() <= 1e-5 * max(1.0, np.abs(r[k][fin]).max()), k


def test_hooks_with_mask_vs_oracle(hh, orc, ctx):
    """The per-hook entry points (process_windows!, correlation planes,
    uncertainty_sweep!, the ensemble's accumulate_planes!) with the
    context's mask: every job's window loaded as load_windows! loads it with
    the mask (src/correlators.jl:199-227), jobs as pass_grid leaves them."""
    n = 160
    A, B, m = disk_scene(hh, n=n, seed=73, speckle=0.01)
    p = hh.PIVParameters(window_size=32, overlap=16, padding=True, apodization="gauss", uncertainty=True)
    gm = orc.grid_mask(p.to_struct(), n, n, m)
    x, y, jobs = hh.pass_grid(p, n, n)
    keep = [k for k in range(jobs.shape[1]) if not gm[jobs[0, k] - 1, jobs[1, k] - 1]]
    jobs = np.asfortranarray(jobs[:, keep])
    ny, nx = y.size, x.size
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    o = orc.process_windows(p.to_struct(), A64, B64, jobs, ny, nx, uncertainty=True, planes=True, mask=m)
    ctx.set_pair(A, B)
    ctx.set_mask(m, 0.5)
    try:
        say("hooks with mask")
        r = ctx.process_windows(p, jobs, ny, nx, src=hh.HH_SRC_RAW, uncertainty=True)
        planes = ctx.correlation_planes(p, jobs)
        uu, uv = ctx.uncertainty_sweep(p, jobs, ny, nx, src=hh.HH_SRC_RAW)
        ctx.ensemble_begin(p, jobs, uncertainty=True)
        for _ in range(2):
            ctx.ensemble_add(hh.HH_SRC_RAW)
        eplanes = ctx.ensemble_planes()
    finally:
        ctx.set_mask(None)
    cells = (jobs[0] - 1, jobs[1] - 1)
    du = np.abs(r["u"][cells] - o["u"][cells])
    dv = np.abs(r["v"][cells] - o["v"][cells])
    say("hook u/v max", du.max(), dv.max())
    assert np.count_nonzero((du > 1e-3) | (dv > 1e-3)) <= 1
    scale = np.abs(o["planes"]).max(axis=(0, 1))
    rel = (np.abs(planes - o["planes"]).max(axis=(0, 1)) / scale).max()
    erel = (np.abs(eplanes - 2 * o["planes"]).max(axis=(0, 1)) / (2 * scale)).max()
    say("hook planes max rel", rel, "ensemble", erel)
    assert rel < 1e-5 and erel < 1e-5
    for got in ((r["unc_u"], r["unc_v"]), (uu, uv)):
        for g, ref in zip(got, (o["unc_u"], o["unc_v"])):
            fin = np.isfinite(ref[cells])
            assert np.array_equal(np.isfinite(g[cells]), fin)
            rr = np.abs(g[cells][fin] - ref[cells][fin]) / np.maximum(np.abs(ref[cells][fin]), 1e-6)
            assert rr.max() < 0.01


@pytest.mark.parametrize("form", ["fft32", "mfma16"])
def test_run_piv_mask_with_search_area_vs_oracle(hh, orc, form):
    """Masks with an enlarged search area (load_windows!, src/correlators.jl:
    230-262: window and search area each with its own valid-pixel mean; the
    grid mask counts both fractions): a 16-px window in a 32-px area on the
    first pass, equal areas with uncertainty on the last."""
    A, B, m = disk_scene(hh, seed=79, speckle=0.01)
    passes = [hh.PIVParameters(window_size=16, search_area_size=32, overlap=8, padding=True, apodization="gauss"),
              hh.PIVParameters(window_size=16, overlap=8, padding=True, apodization="gauss", max_iterations=2,
                               convergence_tol=0.0, uncertainty=True)]
    say("mask + search area", form)
    with hh.Context(0) as c:
        c.set_form(hh.HH_FORM_MFMA16 if form == "mfma16" else hh.HH_FORM_FFT32)
        g, o = run_masked(hh, orc, c, A, B, m, passes)
    check_masked(g, o, uq=True)


def test_single_pass_mask_with_search_area_uq_vs_oracle(hh, orc, ctx):
    # the uncertainty of an enlarged-search pass with a mask reads window
    # tiles of both frames (src/pipeline.jl:984-989)
    A, B, m = disk_scene(hh, seed=83, speckle=0.01)
    p = hh.PIVParameters(window_size=16, search_area_size=32, overlap=8, uncertainty=True)
    g, o = run_masked(hh, orc, ctx, A, B, m, [p])
    check_masked(g, o, max_tol=1e-3, uq=True)


def test_hooks_and_ensemble_mask_with_search_area_vs_oracle(hh, orc, ctx):
    n = 160
    A, B, m = disk_scene(hh, n=n, seed=89, speckle=0.01)
    p = hh.PIVParameters(window_size=16, search_area_size=32, overlap=8, padding=True, apodization="gauss",
                         uncertainty=True)
    gm = orc.grid_mask(p.to_struct(), n, n, m)
    x, y, jobs = hh.pass_grid(p, n, n)
    keep = [k for k in range(jobs.shape[1]) if not gm[jobs[0, k] - 1, jobs[1, k] - 1]]
    jobs = np.asfortranarray(jobs[:, keep])
    ny, nx = y.size, x.size
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    o = orc.process_windows(p.to_struct(), A64, B64, jobs, ny, nx, uncertainty=True, planes=True, mask=m)
    ctx.set_pair(A, B)
    ctx.set_mask(m, 0.5)
    try:
        r = ctx.process_windows(p, jobs, ny, nx, src=hh.HH_SRC_RAW, uncertainty=True)
        planes = ctx.correlation_planes(p, jobs)
        uu, uv = ctx.uncertainty_sweep(p, jobs, ny, nx, src=hh.HH_SRC_RAW)
        e = ctx.run_piv_ensemble([(A, B), (A, B)], [p])
    finally:
        ctx.set_mask(None)
    cells = (jobs[0] - 1, jobs[1] - 1)
    du = np.abs(r["u"][cells] - o["u"][cells])
    dv = np.abs(r["v"][cells] - o["v"][cells])
    scale = np.abs(o["planes"]).max(axis=(0, 1))
    rel = (np.abs(planes - o["planes"]).max(axis=(0, 1)) / scale).max()
    say("mask + search hooks: u/v max", du.max(), dv.max(), "planes rel", rel)
    assert np.count_nonzero((du > 1e-3) | (dv > 1e-3)) <= 1 and rel < 3e-5
    for got in ((r["unc_u"], r["unc_v"]), (uu, uv)):
        for gq, ref in zip(got, (o["unc_u"], o["unc_v"])):
            fin = np.isfinite(ref[cells])
            assert np.array_equal(np.isfinite(gq[cells]), fin)
            rr = np.abs(gq[cells][fin] - ref[cells][fin]) / np.maximum(np.abs(ref[cells][fin]), 1e-6)
            assert rr.max() < 0.01
    # the ensemble of two identical pairs under the same mask
    oe = orc.run_piv_ensemble([(A64, B64), (A64, B64)], [p.to_struct()], mask=m)
    check_masked(e, oe, uq=True)
