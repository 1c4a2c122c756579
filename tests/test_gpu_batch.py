"""Batches of pairs of uneven sizes (run_piv_sequence, src/io.jl:229-247, on
the device) and hipGraph replay of the whole schedule, through the C ABI.

Bars: every pair of a batch gives exactly (bitwise, NaN positions included)
what run_piv gives on that pair alone, since both run the same per-window,
per-pixel and per-cell arithmetic; a graph replay gives exactly the eager
launch sequence's result; and a batch pair against the FP64 oracle meets the
run_piv bars (flags within the near-tie allowance, u/v RMS < 5e-3 px)."""
import re
import sys

import numpy as np
import pytest

from parity_bars import field_parity

pytestmark = pytest.mark.gpu

KEYS = ("u", "v", "peak_ratio", "correlation_moment", "uncertainty_u", "uncertainty_v", "outliers", "x", "y")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same(a, b):
    return all(np.array_equal(a[k], b[k], equal_nan=(k != "outliers")) for k in KEYS)


def diff_report(a, b):
    out = []
    for k in KEYS:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        if x.shape != y.shape:
            out.append(f"{k}: shape {x.shape} vs {y.shape}")
        elif not np.array_equal(x, y, equal_nan=(k != "outliers")):
            d = np.abs(x.astype(float) - y.astype(float))
            out.append(f"{k}: {int(np.sum(~(np.isclose(x, y, rtol=0, atol=0, equal_nan=True))))} cells, "
                       f"max {np.nanmax(d) if np.isfinite(d).any() else 'nan'}")
    return "; ".join(out)


SMALL = ((256, 256), (192, 320), (320, 256), (160, 160), (256, 384))


@pytest.fixture(scope="module")
def small_pairs(hh):
    return hh.synthetic.config_batch(npairs=7, seed0=204000, sizes=SMALL)


def schedule(hh, **final):
    f = dict(overlap=12, max_iterations=2, convergence_tol=0.0, uncertainty=True)
    f.update(final)
    return hh.multipass_parameters([32, 16], padding=True, apodization="gauss", final=f)


@pytest.mark.parametrize("form", ["fft32", "mfma16"])
def test_batch_equals_single_runs(hh, small_pairs, form):
    passes = schedule(hh)
    fm = hh.HH_FORM_MFMA16 if form == "mfma16" else hh.HH_FORM_FFT32
    say("batch of", len(small_pairs), "pairs, form", form)
    with hh.Context(0) as cb, hh.Context(0) as cs:
        cb.set_form(fm)
        cs.set_form(fm)
        cb.set_pairs([(p.A, p.B) for p in small_pairs])
        rb = cb.run_piv_batch(passes)
        for q, p in enumerate(small_pairs):
            cs.set_pair(p.A, p.B)
            rs = cs.run_piv(passes)
            assert same(rb[q], rs), f"pair {q} {p.A.shape}: {diff_report(rb[q], rs)}"
            assert np.isfinite(rb[q]["uncertainty_u"]).mean() > 0.9


def test_batch_dynamic_convergence_per_pair(hh, small_pairs):
    """Iterating final pass with convergence_tol > 0: each pair stops on its
    own q95 field change, exactly as run_piv on it alone."""
    passes = schedule(hh, max_iterations=5, convergence_tol=0.02)
    say("dynamic batch")
    with hh.Context(0) as cb, hh.Context(0) as cs:
        cb.set_pairs([(p.A, p.B) for p in small_pairs])
        rb = cb.run_piv_batch(passes)
        for q, p in enumerate(small_pairs):
            cs.set_pair(p.A, p.B)
            rs = cs.run_piv(passes)
            assert same(rb[q], rs), f"pair {q} {p.A.shape}: {diff_report(rb[q], rs)}"


def test_graph_replays_match_eager(hh, small_pairs):
    """The captured schedule replayed several times, with eager work and
    synchronizations in between: every replay equals the eager result."""
    passes = schedule(hh)
    pairs = [(p.A, p.B) for p in small_pairs[:4]]
    say("graph replay")
    with hh.Context(0) as cg, hh.Context(0) as ce:
        ce.set_graphs(False)
        ce.set_pairs(pairs)
        ref = ce.run_piv_batch(passes)
        assert ce.graph_replays() == 0
        cg.set_pairs(pairs)
        for rep in range(4):
            got = cg.run_piv_batch(passes)
            for q in range(len(pairs)):
                assert same(got[q], ref[q]), f"replay {rep} pair {q}: {diff_report(got[q], ref[q])}"
            # eager work on the same context between replays
            cg.smooth_field(np.ones((8, 8), np.float32))
            cg.synchronize()
        assert cg.graph_replays() == 4
        # a new pair upload into the same atlas is seen by the next replay
        flipped = [(B, A) for A, B in pairs]
        cg.set_pairs(flipped)
        ce.set_pairs(flipped)
        assert all(same(a, b) for a, b in zip(cg.run_piv_batch(passes), ce.run_piv_batch(passes)))


def test_single_pair_graph_matches_eager(hh):
    pair = hh.synthetic.config_vortex_1024(n=384)
    passes = hh.multipass_parameters([64, 32, 16], padding=True, apodization="gauss",
                                     final=dict(overlap=12, max_iterations=3, convergence_tol=0.0,
                                                uncertainty=True))
    with hh.Context(0) as cg, hh.Context(0) as ce:
        ce.set_graphs(False)
        ce.set_pair(pair.A, pair.B)
        cg.set_pair(pair.A, pair.B)
        ref = ce.run_piv(passes)
        for rep in range(3):
            got = cg.run_piv(passes)
            assert same(got, ref), f"replay {rep}: {diff_report(got, ref)}"
        assert cg.graph_replays() == 3


def test_batch_pair_vs_oracle(hh, orc, small_pairs):
    passes = schedule(hh)
    pairs = small_pairs[:3]
    with hh.Context(0) as cb:
        cb.set_pairs([(p.A, p.B) for p in pairs])
        rb = cb.run_piv_batch(passes)
    for q, p in enumerate(pairs):
        o = orc.run_piv(p.A.astype(np.float64), p.B.astype(np.float64), [s.to_struct() for s in passes], uq_pert=True,
                        prec="f64")
        field_parity(rb[q], o, label=f"batch pair {q} {p.A.shape}", rms_tol=5e-3)


def test_run_piv_sequence_api(hh, small_pairs):
    passes = schedule(hh)
    seen = []
    res = hh.run_piv_sequence([(p.A, p.B) for p in small_pairs[:5]], passes, batch=2,
                              on_result=lambda i, r: seen.append((i, r.u.shape)))
    assert [i for i, _ in seen] == [1, 2, 3, 4, 5]
    for p, r in zip(small_pairs[:5], res):
        single = hh.run_piv(p.A, p.B, passes)
        assert np.array_equal(r.u, single.u, equal_nan=True)
        assert np.array_equal(r.outliers, single.outliers)


@pytest.mark.parametrize("form", ["fft32", "mfma16"])
def test_autotune_never_changes_results(hh, small_pairs, form):
    passes = schedule(hh)
    fm = hh.HH_FORM_MFMA16 if form == "mfma16" else hh.HH_FORM_FFT32
    pairs = [(p.A, p.B) for p in small_pairs[:3]]
    with hh.Context(0) as ct, hh.Context(0) as cn:
        cn.set_autotune(False)
        for c in (ct, cn):
            c.set_form(fm)
            c.set_pairs(pairs)
        rt, rn = ct.run_piv_batch(passes), cn.run_piv_batch(passes)
        rep = ct.tune_report()
        say("tuned", rep)
        assert rep and cn.tune_report() == {}
        fams = " ".join(rep)
        assert ("mfma32" in fams) == (form == "mfma16")
        for key, e in rep.items():
            assert len(e["ms"]) >= 2 and all(v > 0 for v in e["ms"].values())
        for q in range(len(pairs)):
            assert same(rt[q], rn[q]), f"pair {q}: {diff_report(rt[q], rn[q])}"


def test_ensemble_resident_rerun_matches(hh):
    """run_piv_ensemble on the pairs the previous call left resident gives
    the same result as the call that uploaded them; a count or size that does
    not match the resident pairs is a call-order error."""
    pairs, _ = hh.synthetic.config_ensemble(npairs=6, nr=160, nc=192)
    p = hh.PIVParameters(window_size=32, overlap=16, padding=True, apodization="gauss")
    with hh.Context(0) as c:
        first = c.run_piv_ensemble([(q.A, q.B) for q in pairs], [p])
        again = c.run_piv_ensemble_resident(len(pairs), pairs[0].A.shape, [p])
        assert same(first, again), diff_report(first, again)
        with pytest.raises(hh.DeviceError):
            c.run_piv_ensemble_resident(len(pairs) + 1, pairs[0].A.shape, [p])


@pytest.mark.parametrize("form", ["fft32", "mfma16"])
def test_autotune_kernel_families(hh, orc, small_pairs, form):
    """Level-2 tuning chooses the kernel family per FFT side as well: every
    pass's report lists the family candidates with their times; a run with
    each pass forced to each of its families meets the parity bars against
    the FP64 oracle; the tuned run equals a run with its choices preloaded."""
    passes = schedule(hh)
    fm = hh.HH_FORM_MFMA16 if form == "mfma16" else hh.HH_FORM_FFT32
    p = small_pairs[0]
    o = orc.run_piv(p.A.astype(np.float64), p.B.astype(np.float64), [s.to_struct() for s in passes],
                    uq_pert=True)
    with hh.Context(0) as c:
        c.set_form(fm)
        c.set_autotune(True, level=2)
        c.set_pair(p.A, p.B)
        tuned = c.run_piv(passes)
        rep = c.tune_report()
    say("level-2 report", rep)
    assert rep
    for key, e in rep.items():
        fams = {n.split()[0] for n in e["ms"]}
        assert all(n.startswith("family=") for n in fams) and len(fams) >= 2, (key, fams)
    with hh.Context(0) as c:
        c.set_form(fm)
        c.tune_load(rep)
        c.set_pair(p.A, p.B)
        again = c.run_piv(passes)
    assert same(tuned, again), diff_report(tuned, again)
    for fam in (0, 1, 2):
        forced = {k: dict(e, choice=dict(e["choice"], family=fam, waves_per_simd=0)) for k, e in rep.items()}
        with hh.Context(0) as c:
            c.set_form(fm)
            c.tune_load(forced)
            c.set_pair(p.A, p.B)
            g = c.run_piv(passes)
        field_parity(g, o, label=f"{form} family {fam}", rms_tol=5e-3)


def test_every_launch_geometry_gives_the_same_field(hh, small_pairs):
    """Every launch-geometry candidate the level-1 tuner times for the
    matrix-core kernels (resident waves per SIMD, which for the 16-px kernel
    also selects its 128-register build at 4 waves; window order) gives the
    untuned field bit for bit: the geometry only changes which wave takes
    which window."""
    passes = schedule(hh)
    p = small_pairs[0]
    with hh.Context(0) as c:
        c.set_form(hh.HH_FORM_MFMA16)
        c.set_autotune(False)
        c.set_pair(p.A, p.B)
        base = c.run_piv(passes)
    with hh.Context(0) as c:
        c.set_form(hh.HH_FORM_MFMA16)
        c.set_pair(p.A, p.B)
        c.run_piv(passes)
        rep = c.tune_report()
    assert any(k.startswith("mfma32 ") for k in rep), rep
    nrun = 0
    for key, e in rep.items():
        for name in e["ms"]:
            m = re.search(r"waves_per_simd=(\d+)", name)
            if not m:
                continue
            w, order = int(m.group(1)), int("order=linear" in name)
            forced = dict(rep)
            forced[key] = dict(e, choice=dict(e["choice"], waves_per_simd=w, linear_order=order))
            with hh.Context(0) as c:
                c.set_form(hh.HH_FORM_MFMA16)
                c.tune_load(forced)
                c.set_pair(p.A, p.B)
                g = c.run_piv(passes)
            say(f"{key}: {name}")
            assert same(base, g), f"{key} {name}: {diff_report(base, g)}"
            nrun += 1
    assert nrun >= 8
