"""The measured workloads against the FP64 oracle, at their full sizes.

* The 4 MP solid-body rotation (2048 x 2048 Float32, 0.03 ppp, 64/32 ->
  32/16 -> 16/12 x 3 sweeps, padded Gaussian passes, uncertainty on the final
  pass): the workload the throughput figure is quoted on, in both correlation
  forms (FP32 FFT and split-FP16 matrix-core DFT).
* The 1024 x 1024 Lamb-Oseen vortex with the same schedule.
* The reference's default early exit (convergence_tol = 0.05,
  src/pipeline.jl:779-788,849-866): the device stops after the same sweep as
  the oracle.
* replace_outliers = false on an iterating pass (keep_measured,
  src/pipeline.jl:720-723,768-770,826-829): flagged cells hold the measured
  values of the last sweep.

Bars (tests/parity_bars.py): flags identical, NaN/Inf classes identical,
|du|, |dv| <= 1e-3 px at every cell, RMS <= 0.005 px, peak ratio and moment
within 1 % relative, sigma within 1 % at every finite cell where the FP64
sigma is well-conditioned (the others are counted and reported).
"""
import numpy as np
import pytest

from parity_bars import field_parity, field_stats, say

pytestmark = pytest.mark.gpu

FORMS = ("fft32", "mfma16")


def schedule(hh, **final):
    f = dict(overlap=12, max_iterations=3, convergence_tol=0.0, uncertainty=True)
    f.update(final)
    return hh.multipass_parameters([64, 32, 16], padding=True, apodization="gauss", final=f)


def run_device(hh, form, A, B, passes):
    with hh.Context(0) as c:
        c.set_form(hh.HH_FORM_MFMA16 if form == "mfma16" else hh.HH_FORM_FFT32)
        c.set_pair(A, B)
        return c.run_piv(passes)


def run_oracle(orc, A, B, passes):
    return orc.run_piv(A.astype(np.float64), B.astype(np.float64), [p.to_struct() for p in passes], prec="f64",
                       uq_pert=True, nthreads=orc.default_threads())


@pytest.fixture(scope="module")
def rotation(hh, orc):
    pair = hh.synthetic.config_rotation_4mp()
    passes = schedule(hh)
    say("oracle: 2048^2 rotation, FP64")
    return pair, passes, run_oracle(orc, pair.A, pair.B, passes)


@pytest.fixture(scope="module")
def vortex(hh, orc):
    pair = hh.synthetic.config_vortex_1024()
    passes = schedule(hh)
    say("oracle: 1024^2 vortex, FP64")
    return pair, passes, run_oracle(orc, pair.A, pair.B, passes)


@pytest.mark.parametrize("form", FORMS)
def test_rotation_4mp_vs_fp64_oracle(hh, rotation, form):
    pair, passes, o = rotation
    say("device: 2048^2 rotation", form)
    g = run_device(hh, form, pair.A, pair.B, passes)
    st = field_parity(g, o, label=f"rotation 4mp {form}", rms_tol=0.005)
    say("rms vs midpoint truth: device", hh.synthetic.rms_error(g, pair.flow), "oracle",
        hh.synthetic.rms_error(o, pair.flow))
    assert st["cells"] == 509 * 509


@pytest.mark.parametrize("form", FORMS)
def test_vortex_1024_vs_fp64_oracle(hh, vortex, form):
    pair, passes, o = vortex
    say("device: 1024^2 vortex", form)
    g = run_device(hh, form, pair.A, pair.B, passes)
    field_parity(g, o, label=f"vortex 1024 {form}", rms_tol=0.005)


def _same(a, b):
    return all(np.array_equal(np.asarray(a[k]), np.asarray(b[k]), equal_nan=True)
               for k in ("u", "v", "peak_ratio", "correlation_moment", "outliers"))


@pytest.mark.parametrize("form", FORMS)
def test_early_exit_stops_with_the_oracle(hh, orc, form):
    # the final pass iterates until the q95 field change drops below 0.05 px
    # (at most 5 sweeps); the sweep it stopped after is found as the fixed
    # sweep count whose result is identical, on each side
    pair = hh.synthetic.config_vortex_1024(n=512)
    tol = schedule(hh, max_iterations=5, convergence_tol=0.05, uncertainty=False)
    g = run_device(hh, form, pair.A, pair.B, tol)
    o = run_oracle(orc, pair.A, pair.B, tol)
    field_parity(g, o, label=f"early exit {form}", rms_tol=0.005, uncertainty=False)
    kg = ko = None
    for k in range(1, 6):
        fixed = schedule(hh, max_iterations=k, convergence_tol=0.0, uncertainty=False)
        if kg is None and _same(run_device(hh, form, pair.A, pair.B, fixed), g):
            kg = k
        if ko is None and _same(run_oracle(orc, pair.A, pair.B, fixed), o):
            ko = k
    say("early exit", form, "device stopped after sweep", kg, "oracle after", ko)
    assert ko is not None and ko < 5, "the workload should converge before the sweep cap"
    assert kg == ko


@pytest.mark.parametrize("form", FORMS)
def test_keep_measured_vs_oracle(hh, orc, form):
    # sweeps replace outliers to build the next predictor; the returned field
    # restores the measured values at cells still flagged after the last sweep
    pair = hh.synthetic.config_vortex_1024(n=512)
    passes = hh.multipass_parameters([64, 32, 16], padding=True, apodization="gauss",
                                     final=dict(overlap=12, max_iterations=3, convergence_tol=0.0,
                                                replace_outliers=False))
    g = run_device(hh, form, pair.A, pair.B, passes)
    o = run_oracle(orc, pair.A, pair.B, passes)
    st = field_parity(g, o, label=f"keep_measured {form}", rms_tol=0.005, uncertainty=False)
    assert st["outliers"] > 0, "the scene should leave flagged cells"
    rep = run_device(hh, form, pair.A, pair.B, hh.multipass_parameters(
        [64, 32, 16], padding=True, apodization="gauss",
        final=dict(overlap=12, max_iterations=3, convergence_tol=0.0, replace_outliers=True)))
    flagged = o["outliers"]
    # unflagged cells are the same either way; flagged ones differ from the
    # replaced values (the measured vector is kept)
    assert np.array_equal(rep["u"][~flagged], g["u"][~flagged])
    assert np.any(rep["u"][flagged] != g["u"][flagged]) or np.any(rep["v"][flagged] != g["v"][flagged])
    say("keep_measured", form, field_stats(g, o, cells=flagged))
