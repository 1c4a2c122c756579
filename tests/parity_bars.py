"""Field-level parity bars shared by the GPU tests: a device PIVResult (dict of
grids) against the FP64 oracle's on the same inputs.

Bars (the acceptance criteria of the device path):
  * outlier flags identical cell for cell;
  * NaN / Inf classes identical in every field (peak_ratio is Inf where only
    one peak was found, NaN where none was positive);
  * u, v within ``tol_px`` (1e-3 px) of the oracle at every finite cell;
  * RMS of the u/v deviation below ``rms_tol`` when given (0.005 px: the
    split-FP16 form's bar);
  * peak ratio and correlation moment within ``quality_tol`` relative at every
    finite cell (both come from the FP32 plane, so their error follows the
    plane's relative error);
  * sigma_u, sigma_v within ``sigma_tol`` (1 %) relative, or within
    ``sigma_abs`` (1e-5 px, a hundredth of the u/v bar, for sigma near 0), at
    every finite cell where the FP64 sigma is well-conditioned. Sigma is a threshold-laden function of 44 FP64
    statistics (a ring is summed only while its maximum reaches 0.05 S00;
    sigma_dC^2 is clipped at 0; log(lo) needs lo > 0; sigma_dC^2 is a
    difference of nearly equal sums), so at some cells float32-level error
    in the images alone moves it by more than 1 %. The oracle reports its
    sigma recomputed from its own final images rounded to float32 and scaled
    by 1 + 2^-21 h per pixel, with the predictor moved by up to 1e-5 px
    (``uq_pert_u/v``), and its FP64 statistics (``uq_stats``). A cell where
    the perturbation moves sigma by more than 0.5 % (or changes its finite
    class), or whose ring truncation decides within 1e-4 S00 of its
    threshold, is ill-conditioned: counted and reported, and exempt from the
    1 % bar.
  * FP64 near-ties are reported separately: a window whose primary peak is
    within 1e-5 relative of a neighbouring plane value or of the secondary
    peak (the oracle lists them for every pass and sweep as ``ties``) may
    legitimately resolve the other way in float32, and the difference then
    travels through the predictor into the later passes. Cells within
    ``tie_radius`` (128 px) of such a window are counted and reported and
    exempt from the bars above; everything else must meet them.
Every statistic is printed before any assert so a failing run still reports
all of them.
"""
import math
import sys

import numpy as np


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rel(g, o, floor):
    return np.abs(g - o) / np.maximum(np.abs(o), floor)


def ring_margins(stats, rings):
    """Per cell and component, the smallest distance of a ring-truncation
    test of the Wieneke finalization from its threshold, in units of S00:
    min over the rings it evaluates (up to and including the ring that stops
    the sum) of |max(ring) - 0.05 S00| / S00 (finalize_uncertainty,
    src/uncertainty.jl:163-168). stats: (44, 2, ny, nx)."""
    out = []
    for comp in range(2):
        s = stats[:, comp]
        S00 = s[3]
        with np.errstate(all="ignore"):
            marg = np.full(S00.shape, np.inf)
            live = np.ones(S00.shape, bool)
            for r in range(len(rings) - 1):
                mx = s[4 + rings[r]:4 + rings[r + 1]].max(axis=0)
                d = np.abs(mx - 0.05 * S00) / np.abs(S00)
                marg = np.where(live, np.minimum(marg, d), marg)
                live &= ~(mx < 0.05 * S00)
        out.append(marg)
    return out


def sigma_ill_conditioned(o, rel=5e-3, margin=1e-4, rings=None):
    """Bool grids (u, v): cells whose FP64 sigma moves by more than ``rel``
    relative (or changes finite class) under the oracle's float32-level
    perturbation of its final images, or whose ring truncation decides
    within ``margin`` S00 of its threshold."""
    if rings is None:
        import oracle

        rings = oracle.uq_offsets()[1]
    margins = ring_margins(o["uq_stats"], rings)
    out = []
    for comp, (k, kp) in enumerate((("uncertainty_u", "uq_pert_u"), ("uncertainty_v", "uq_pert_v"))):
        a, b = np.asarray(o[k], np.float64), np.asarray(o[kp], np.float64)
        fa, fb = np.isfinite(a), np.isfinite(b)
        with np.errstate(all="ignore"):
            moved = fa & fb & (np.abs(b - a) > rel * np.maximum(np.abs(a), 1e-6))
        out.append(moved | (fa != fb) | (margins[comp] < margin))
    return out


def tie_region(o, radius=128.0):
    """Final-grid cells within ``radius`` px (Chebyshev) of a near-tie window
    of any pass or sweep of the oracle run."""
    X, Y = np.meshgrid(np.asarray(o["x"], np.float64), np.asarray(o["y"], np.float64))
    near = np.zeros(X.shape, bool)
    for _, _, ty, tx in np.asarray(o.get("ties", np.zeros((0, 4)))):
        near |= (np.abs(Y - ty) <= radius) & (np.abs(X - tx) <= radius)
    return near


def field_stats(g, o, cells=None):
    """Deviation statistics of device result g from oracle result o.
    ``cells`` (bool grid, optional) restricts every statistic."""
    sel = np.ones(o["u"].shape, bool) if cells is None else np.asarray(cells, bool)
    st = {"cells": int(sel.sum())}
    st["flag_diffs"] = int(((g["outliers"] != o["outliers"]) & sel).sum())
    st["outliers"] = int((o["outliers"] & sel).sum())
    cls = {}
    for k in ("u", "v", "peak_ratio", "correlation_moment", "uncertainty_u", "uncertainty_v"):
        if k not in g or k not in o:
            continue
        a, b = np.asarray(g[k], np.float64), np.asarray(o[k], np.float64)
        cls[k] = int(((np.isnan(a) != np.isnan(b)) | (np.isinf(a) != np.isinf(b)))[sel].sum())
    st["class_diffs"] = cls
    fin = sel & np.isfinite(g["u"]) & np.isfinite(o["u"]) & np.isfinite(g["v"]) & np.isfinite(o["v"])
    du = np.abs(g["u"].astype(np.float64) - o["u"])[fin]
    dv = np.abs(g["v"].astype(np.float64) - o["v"])[fin]
    d = np.concatenate([du, dv])
    st["max_du"] = float(du.max(initial=0.0))
    st["max_dv"] = float(dv.max(initial=0.0))
    st["rms_px"] = float(math.sqrt(np.mean(d ** 2))) if d.size else 0.0
    for k, floor in (("peak_ratio", 1e-30), ("correlation_moment", 1e-30)):
        a, b = np.asarray(g[k], np.float64), np.asarray(o[k], np.float64)
        f = sel & np.isfinite(a) & np.isfinite(b)
        r = _rel(a[f], b[f], floor)
        st[k + "_rel_max"] = float(r.max(initial=0.0))
        st[k + "_rel_median"] = float(np.median(r)) if r.size else 0.0
    for k in ("uncertainty_u", "uncertainty_v"):
        if k not in g:
            continue
        a, b = np.asarray(g[k], np.float64), np.asarray(o[k], np.float64)
        f = sel & np.isfinite(a) & np.isfinite(b)
        r = _rel(a[f], b[f], 1e-6)
        st[k + "_rel_max"] = float(r.max(initial=0.0))
        st[k + "_rel_median"] = float(np.median(r)) if r.size else 0.0
        st[k + "_finite"] = int(f.sum())
    return st


def field_parity(g, o, label="", tol_px=1e-3, rms_tol=None, quality_tol=1e-2, sigma_tol=1e-2, uncertainty=True,
                 cells=None, sigma_abs=1e-5, tie_radius=128.0):
    """Assert the parity bars; returns the statistics."""
    if "ties" in o and len(o["ties"]):
        near = tie_region(o, tie_radius)
        base = np.ones(o["u"].shape, bool) if cells is None else np.asarray(cells, bool)
        inside = field_stats(g, o, base & near)
        say("near-tie region", label, {"ties": int(o.get("ties_total", len(o["ties"]))),
                                       "cells": inside["cells"], "flag_diffs": inside["flag_diffs"],
                                       "max_du": inside["max_du"], "max_dv": inside["max_dv"]})
        cells = base & ~near
    st = field_stats(g, o, cells)
    ill = None
    if uncertainty and "uq_pert_u" in o:
        ill = sigma_ill_conditioned(o)
        sel = np.ones(o["u"].shape, bool) if cells is None else np.asarray(cells, bool)
        for comp, k in enumerate(("uncertainty_u", "uncertainty_v")):
            a, b = np.asarray(g[k], np.float64), np.asarray(o[k], np.float64)
            f = sel & np.isfinite(a) & np.isfinite(b)
            over = f & (_rel(a, b, 1e-6) > sigma_tol) & (np.abs(a - b) > sigma_abs)
            st[k + "_ill_conditioned"] = int((ill[comp] & sel).sum())
            st[k + "_over_bar"] = int(over.sum())
            st[k + "_over_bar_well_conditioned"] = int((over & ~ill[comp]).sum())
            for i, j in np.argwhere(over & ~ill[comp])[:5]:
                say("  sigma over the bar at", k, (int(i), int(j)), "oracle", float(b[i, j]), "device",
                    float(a[i, j]), "perturbed", float(o["uq_pert_" + k[-1]][i, j]), "stats",
                    o["uq_stats"][:8, comp, i, j].tolist())
            good = f & ~ill[comp]
            r = _rel(a[good], b[good], 1e-6)
            st[k + "_rel_max_well_conditioned"] = float(r.max(initial=0.0))
            cd = sel & ((np.isnan(a) != np.isnan(b)) | (np.isinf(a) != np.isinf(b)))
            st["class_diffs"][k] = int((cd & ~ill[comp]).sum())
            st[k + "_class_diffs_ill_conditioned"] = int((cd & ill[comp]).sum())
    say("parity", label, st)
    assert st["flag_diffs"] == 0, (label, "outlier flags differ", st)
    keys = ["u", "v", "peak_ratio", "correlation_moment"] + (["uncertainty_u", "uncertainty_v"] if uncertainty else [])
    for k in keys:
        assert st["class_diffs"].get(k, 0) == 0, (label, k, "NaN/Inf classes differ", st)
    assert st["max_du"] <= tol_px and st["max_dv"] <= tol_px, (label, "u/v deviation", st)
    if rms_tol is not None:
        assert st["rms_px"] <= rms_tol, (label, "RMS deviation", st)
    assert st["peak_ratio_rel_max"] <= quality_tol, (label, "peak ratio", st)
    assert st["correlation_moment_rel_max"] <= quality_tol, (label, "correlation moment", st)
    if uncertainty:
        for k in ("uncertainty_u", "uncertainty_v"):
            if ill is not None:
                assert st[k + "_over_bar_well_conditioned"] == 0, (label, k, st)
            elif k + "_rel_max" in st:
                assert st[k + "_rel_max"] <= sigma_tol, (label, k, st)
    return st
