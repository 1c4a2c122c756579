"""The field-level parity bars (tests/parity_bars.py) pinned on the CPU: the
oracle's own Float32 restatement of run_piv meets them against the FP64
oracle, and a planted mistake of each kind (one flipped flag, one vector off
by 2e-3 px, one well-conditioned sigma off by 2 %, a ratio off by 2 %, a NaN
class changed) fails them."""
import numpy as np
import pytest

from parity_bars import field_parity, sigma_ill_conditioned


@pytest.fixture(scope="module")
def runs(hh, orc):
    pair = hh.synthetic.config_vortex_1024(n=256)
    passes = hh.multipass_parameters([32, 16], padding=True, apodization="gauss",
                                     final=dict(overlap=8, max_iterations=2, convergence_tol=0.0, uncertainty=True))
    st = [p.to_struct() for p in passes]
    o = orc.run_piv(pair.A.astype(np.float64), pair.B.astype(np.float64), st, uq_pert=True)
    o32 = orc.run_piv(pair.A, pair.B, st, prec="f32")
    return o, o32


def test_float32_restatement_meets_the_bars(orc, runs):
    o, o32 = runs
    st = field_parity(o32, o, label="oracle f32 vs f64", rms_tol=0.005)
    assert st["uncertainty_u_finite"] > 900
    # the conditioning screen is narrow: it removes well under 2 % of cells
    assert st["uncertainty_u_ill_conditioned"] + st["uncertainty_v_ill_conditioned"] < 0.02 * 2 * o["u"].size


def test_perturbed_sigma_is_close(orc, runs):
    # the float32-level perturbation moves sigma by ~1e-5 typically; only a
    # small fraction of cells is ill-conditioned
    o, _ = runs
    for k in ("u", "v"):
        a, b = o["uncertainty_" + k], o["uq_pert_" + k]
        f = np.isfinite(a) & np.isfinite(b)
        r = np.abs(a[f] - b[f]) / np.maximum(np.abs(a[f]), 1e-6)
        assert np.median(r) < 1e-4 and f.sum() > 900


def _planted(o32, key, cell, fn):
    g = {k: (v.copy() if hasattr(v, "copy") else v) for k, v in o32.items()}
    g[key][cell] = fn(g[key][cell])
    return g


def test_planted_mistakes_fail(orc, runs):
    o, o32 = runs
    ill = sigma_ill_conditioned(o)
    fin = np.isfinite(o["uncertainty_u"]) & ~ill[0] & (o["uncertainty_u"] > 1e-3)
    cell = tuple(np.argwhere(fin)[0])
    pr_cell = tuple(np.argwhere(np.isfinite(o["peak_ratio"]))[0])
    cases = [("outliers", cell, lambda x: not x), ("u", cell, lambda x: x + 2e-3),
             ("uncertainty_u", cell, lambda x: x * 1.02), ("peak_ratio", pr_cell, lambda x: x * 1.02),
             ("correlation_moment", cell, lambda x: np.nan)]
    for key, c, fn in cases:
        with pytest.raises(AssertionError):
            field_parity(_planted(o32, key, c, fn), o, label=f"planted {key}", rms_tol=0.005)


def test_ring_margin_marks_the_truncation_edge(orc):
    # a cell whose first ring maximum sits 1e-7 S00 above 0.05 S00: a nudge
    # of the ring below the threshold changes sigma discontinuously, and the
    # margin screen marks it; a cell far from every threshold is not marked
    from parity_bars import ring_margins

    rings = orc.uq_offsets()[1]
    s = np.zeros((44, 2, 1, 2), order="F")
    for cell, top in ((0, 0.05 + 1e-7), (1, 0.2)):
        s[0:4, :, 0, cell] = np.array([1.0, 0.6, 0.58, 1e-3])[:, None]
        s[4:4 + rings[1], :, 0, cell] = top * 1e-3
        s[4 + rings[1]:, :, 0, cell] = -0.01e-3
    m = ring_margins(s, rings)
    assert m[0][0, 0] < 1e-6 and m[0][0, 1] > 0.05
    base = orc.uq_finalize(s[:, 0, 0, 0])
    nudged = s[:, 0, 0, 0].copy()
    nudged[4:4 + rings[1]] = (0.05 - 1e-7) * 1e-3
    assert abs(orc.uq_finalize(nudged) - base) > 0.05 * abs(base)


def test_near_tie_listed_with_its_plateau(hh, orc):
    # batch pair 1 (768 x 1024 shear): in the 32/16 pass the window centred
    # at (752.5, 656.5) has two adjacent plane values 3e-7 apart at its
    # peak; the oracle lists exactly that window as a near-tie
    p = hh.synthetic.config_batch(npairs=2)[1]
    passes = hh.multipass_parameters([64, 32], padding=True, apodization="gauss")
    o = orc.run_piv(p.A.astype(np.float64), p.B.astype(np.float64), [q.to_struct() for q in passes])
    assert o["ties_total"] == 1
    k, it, y, x = o["ties"][0]
    assert (k, it, y, x) == (1, 1, 752.5, 656.5)
    from parity_bars import tie_region

    near = tie_region(o, 32.0)
    assert near.sum() == 5 * 3  # rows 44..46 (y 720.5..752.5) x cols 38..42 of the 47 x 63 grid
