"""The batching and ensemble workloads at their full sizes.

* 64 pairs cycling 512^2, 768 x 1024, 1024 x 1280 and 2048^2 (vortex and
  shear <= 10 px), the 64/32 -> 32/16 -> 16/12 x 3 schedule with uncertainty
  in one device atlas: every pair bitwise equal to run_piv on it alone (both
  forms), one pair of each size against the FP64 oracle at the parity bars.
* The ensemble over 500 pairs of 1024 x 1280 (0.005 ppp, noise 0.05, one
  32/16 padded Gaussian pass, FP64 plane sums): the sums over all 500 pairs
  equal the sums over the first 250 plus those over the last 250 (the
  accumulation is linear, to FP64 rounding), and 24 of the pairs at full
  size match the FP64 oracle's ensemble.
"""
import numpy as np
import pytest

from parity_bars import field_parity, say

pytestmark = pytest.mark.gpu

KEYS = ("u", "v", "peak_ratio", "correlation_moment", "uncertainty_u", "uncertainty_v", "outliers", "x", "y")


def schedule(hh):
    return hh.multipass_parameters([64, 32, 16], padding=True, apodization="gauss",
                                   final=dict(overlap=12, max_iterations=3, convergence_tol=0.0, uncertainty=True))


@pytest.fixture(scope="module")
def batch64(hh):
    return hh.synthetic.config_batch(npairs=64)


@pytest.mark.parametrize("form", ["fft32", "mfma16"])
def test_batch_of_64_equals_single_runs(hh, batch64, form):
    passes = schedule(hh)
    fm = hh.HH_FORM_MFMA16 if form == "mfma16" else hh.HH_FORM_FFT32
    say("batch of 64 full-size pairs", form)
    with hh.Context(0) as cb, hh.Context(0) as cs:
        cb.set_form(fm)
        cs.set_form(fm)
        cb.set_pairs([(p.A, p.B) for p in batch64])
        rb = cb.run_piv_batch(passes)
        for q, p in enumerate(batch64):
            cs.set_pair(p.A, p.B)
            rs = cs.run_piv(passes)
            for k in KEYS:
                assert np.array_equal(rb[q][k], rs[k], equal_nan=(k != "outliers")), (q, p.A.shape, k)


def test_batch_pairs_of_each_size_vs_oracle(hh, orc, batch64):
    passes = schedule(hh)
    with hh.Context(0) as cb:
        cb.set_form(hh.HH_FORM_MFMA16)
        cb.set_pairs([(p.A, p.B) for p in batch64])
        rb = cb.run_piv_batch(passes)
    for q in range(4):  # sizes cycle, so pairs 0..3 cover every size
        p = batch64[q]
        say("oracle for batch pair", q, p.A.shape)
        o = orc.run_piv(p.A.astype(np.float64), p.B.astype(np.float64), [s.to_struct() for s in passes],
                        nthreads=orc.default_threads(), uq_pert=True)
        field_parity(rb[q], o, label=f"batch64 pair {q} {p.A.shape}", rms_tol=0.005)


@pytest.fixture(scope="module")
def ensemble500(hh):
    pairs, flow = hh.synthetic.config_ensemble(npairs=500)
    return [(q.A, q.B) for q in pairs], flow


def test_ensemble_500_sums_are_linear(hh, ensemble500):
    pairs, _ = ensemble500
    p = hh.PIVParameters(window_size=32, overlap=16, padding=True, apodization="gauss")
    nr, nc = pairs[0][0].shape
    x, y, jobs = hh.pass_grid(p, nr, nc)
    sums = []
    with hh.Context(0) as c:
        for part in (pairs, pairs[:250], pairs[250:]):
            c.ensemble_begin(p, jobs)
            for A, B in part:
                c.set_pair(A, B)
                c.ensemble_add(hh.HH_SRC_RAW)
            sums.append(c.ensemble_planes())
    full, a, b = sums
    scale = np.abs(full).max(axis=(0, 1))
    rel = (np.abs(full - (a + b)).max(axis=(0, 1)) / scale).max()
    say("500-pair plane sums vs the halves: max rel", rel)
    assert rel < 1e-12


def test_ensemble_full_size_vs_oracle(hh, orc, ensemble500):
    pairs, flow = ensemble500
    sub = pairs[:24]
    p = hh.PIVParameters(window_size=32, overlap=16, padding=True, apodization="gauss")
    with hh.Context(0) as c:
        g = c.run_piv_ensemble(sub, [p])
    o = orc.run_piv_ensemble([(A.astype(np.float64), B.astype(np.float64)) for A, B in sub], [p.to_struct()],
                             prec="f64", nthreads=orc.default_threads())
    field_parity(g, o, label="ensemble 24 x 1024x1280", uncertainty=False)
    say("rms vs truth", hh.synthetic.rms_error(g, flow, midpoint=False))
