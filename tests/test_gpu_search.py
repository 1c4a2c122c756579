"""Enlarged search areas (search_area_size > window_size) on the device
against the oracle, through the C ABI: the general path of load_windows!
(src/correlators.jl:230-262) and restrict_search_lags! (:347-355) in the
generic correlation kernels, radix-2 for FFT sides 16/32/64/128 and a
direct DFT for 24/48/96 (also equal-area windows of 24 or 48 px).

Bars: planes within 3e-5 of their maximum (FP32 transforms; the unapodized
search area of a small apodized window is the least favourable case, about
1.3e-5 for an 8-px window in a 24-px area); u, v within 1e-3 px per window
(integer lags exact with subpixel :none); run_piv's integer medians of the
reference's search-area scene (test/runtests.jl:66-80); multi-pass and
ensemble runs within the run_piv bars (u/v RMS < 5e-3 px).
"""
import sys

import numpy as np
import pytest

from parity_bars import field_parity

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def shifted_scene(n=128, seed=71, dv=3, du=10):
    # test/runtests.jl:66-69 with numpy's RNG: B is A moved by (dv, du) px
    rng = np.random.default_rng(seed)
    A = rng.random((n, n)).astype(np.float32)
    B = np.zeros((n, n), np.float32)
    B[dv:, du:] = A[:-dv, :-du]
    return np.asfortranarray(A), np.asfortranarray(B)


@pytest.mark.parametrize("w,s,padding,apod", [(16, 32, False, "gauss"), (16, 32, True, "none"),
                                              (32, 64, False, "gauss"), (8, 16, True, "gauss"),
                                              (16, 48, False, "none"), (16, 24, True, "gauss"),
                                              (24, 24, True, "gauss"), (8, 24, False, "gauss"),
                                              (32, 48, True, "none")])
def test_search_planes_and_windows_vs_oracle(hh, orc, ctx, w, s, padding, apod):
    pair = hh.synthetic.particle_pair(160, 160, hh.synthetic.uniform_flow(2.6, -1.3), seed=w + s, noise=0.02)
    p = hh.PIVParameters(window_size=w, search_area_size=s, overlap=w // 2, padding=padding, apodization=apod)
    x, y, jobs = hh.pass_grid(p, 160, 160)
    A64, B64 = pair.A.astype(np.float64), pair.B.astype(np.float64)
    o = orc.process_windows(p.to_struct(), A64, B64, jobs, y.size, x.size, planes=True)
    ctx.set_pair(pair.A, pair.B)
    say("search", w, s, padding, apod, "jobs", jobs.shape[1])
    planes = ctx.correlation_planes(p, jobs)
    M = 2 * s if padding else s
    assert planes.shape == (M, M, jobs.shape[1])
    scale = np.abs(o["planes"]).max(axis=(0, 1))
    rel = (np.abs(planes - o["planes"]).max(axis=(0, 1)) / scale).max()
    r = ctx.process_windows(p, jobs, y.size, x.size, src=hh.HH_SRC_RAW)
    cells = (jobs[0] - 1, jobs[1] - 1)
    du = np.abs(r["u"][cells] - o["u"][cells])
    dv = np.abs(r["v"][cells] - o["v"][cells])
    say("plane rel", rel, "u/v max", du.max(), dv.max())
    assert rel < 3e-5
    assert np.count_nonzero((du > 1e-3) | (dv > 1e-3)) <= 1


@pytest.mark.parametrize("s", [48, 64])
def test_run_piv_search_area_scene(hh, orc, ctx, s):
    # the reference's scene (test/runtests.jl:66-80: a 48-px search area, a
    # direct-DFT side on the device; 64 takes the radix-2 kernel): integer
    # medians (10, 3)
    A, B = shifted_scene()
    p = hh.PIVParameters(window_size=16, search_area_size=s, overlap=8, subpixel_method="none", n_peaks=1,
                         uod_enable=False)
    ctx.set_pair(A, B)
    g = ctx.run_piv([p])
    o = orc.run_piv(A.astype(np.float64), B.astype(np.float64), [p.to_struct()])
    assert np.array_equal(g["x"], o["x"]) and np.array_equal(g["y"], o["y"])
    assert np.median(g["u"]) == 10 and np.median(g["v"]) == 3
    assert np.array_equal(g["u"], o["u"], equal_nan=True) and np.array_equal(g["v"], o["v"], equal_nan=True)
    # test/runtests.jl:98-102: a search-only masked fringe drops the node
    # whose window is clean
    mask = np.zeros_like(A, dtype=bool)
    mask[:16, :48] = True
    ctx.set_mask(mask, 0.3)
    try:
        gm = ctx.run_piv([p])
    finally:
        ctx.set_mask(None)
    om = orc.run_piv(A.astype(np.float64), B.astype(np.float64), [p.to_struct()], mask=mask, mask_threshold=0.3)
    assert np.array_equal(gm["mask"], om["mask"])
    if s == 48:  # 16 x 48 of the 48 x 48 search area: a third, above 0.3
        assert gm["mask"][0, 0]
    assert np.array_equal(gm["u"], om["u"], equal_nan=True) and np.array_equal(gm["v"], om["v"], equal_nan=True)


def test_run_piv_search_multipass_vs_oracle(hh, orc, ctx):
    pair = hh.synthetic.config_vortex_1024(n=256)
    passes = [hh.PIVParameters(window_size=32, search_area_size=64, overlap=16, padding=True, apodization="gauss"),
              hh.PIVParameters(window_size=16, overlap=8, padding=True, apodization="gauss", max_iterations=2,
                               convergence_tol=0.0, uncertainty=True)]
    ctx.set_pair(pair.A, pair.B)
    g = ctx.run_piv(passes)
    o = orc.run_piv(pair.A.astype(np.float64), pair.B.astype(np.float64), [p.to_struct() for p in passes], uq_pert=True)
    field_parity(g, o, label="search multipass", rms_tol=5e-3)


def test_ensemble_search_area_vs_oracle(hh, orc, ctx):
    from conftest import low_snr_pairs

    pairs = [(A.astype(np.float32), B.astype(np.float32)) for A, B in low_snr_pairs(n=96, npairs=4, seed=9)]
    p = hh.PIVParameters(window_size=16, search_area_size=32, overlap=8)
    e = ctx.run_piv_ensemble(pairs, [p])
    o = orc.run_piv_ensemble(pairs, [p.to_struct()], prec="f32")
    keep = ~o["outliers"] & ~e["outliers"]
    assert np.abs(e["u"] - o["u"])[keep].max() < 1e-3 and np.abs(e["v"] - o["v"])[keep].max() < 1e-3
    assert int((e["outliers"] != o["outliers"]).sum()) <= 2
