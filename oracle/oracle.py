"""ctypes binding of the CPU oracle (liboracle.so) — TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module, as the checker. The product library never
loads it. Arrays are numpy, column-major (Fortran order), like the
reference's Julia arrays; positions are 1-based.

``prec`` selects the restatement: ``"f64"`` is the FP64 oracle; ``"f32"``
restates the reference's Float32 pipeline with its Float64 islands.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_lib = None


def build(force: bool = False) -> pathlib.Path:
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "-B" if force else "all"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
    return _lib


def _dt(prec):
    return np.float64 if prec == "f64" else np.float32


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f(a, dtype):
    return np.asfortranarray(a, dtype=dtype)


def _fn(name, prec):
    return getattr(lib(), f"{name}_{prec}")


# ---------------------------------------------------------------------------
# per-function entry points


def apodization(wr, wc, gauss=True, prec="f64"):
    out = np.zeros((wr, wc), dtype=_dt(prec), order="F")
    _fn("orc_apodization", prec)(C.c_int(wr), C.c_int(wc), C.c_int(1 if gauss else 0), _p(out))
    return out


def _plane_shape(pass_struct):
    """FFT side of a pass: the search area, doubled when padded."""
    sr, sc = pass_struct.search_area_size[0], pass_struct.search_area_size[1]
    return (2 * sr, 2 * sc) if pass_struct.padding else (sr, sc)


def gain(pass_struct, prec="f64"):
    m1, m2 = _plane_shape(pass_struct)
    out = np.zeros((m1, m2), dtype=_dt(prec), order="F")
    _fn("orc_gain", prec)(C.byref(pass_struct), _p(out))
    return out


def correlation_plane(pass_struct, subA, subB, prec="f64"):
    dt = _dt(prec)
    m1, m2 = _plane_shape(pass_struct)
    a, b = _f(subA, dt), _f(subB, dt)
    R = np.zeros((m1, m2), dtype=dt, order="F")
    _fn("orc_correlation_plane", prec)(C.byref(pass_struct), _p(a), _p(b), _p(R))
    return R


def is_local_maximum(R, r, c, prec="f64"):
    R = _f(R, _dt(prec))
    return bool(_fn("orc_is_local_maximum", prec)(_p(R), C.c_int(R.shape[0]), C.c_int(R.shape[1]),
                                                  C.c_int(r), C.c_int(c)))


def find_peaks(R, k, peak_finder="regionalmax", exclusion_radius=2, prec="f64"):
    """find_peaks(R, k): list of (value, (row, col)) 1-based."""
    dt = _dt(prec)
    R = _f(R, dt)
    vals = np.zeros(k, dtype=dt)
    locs = np.zeros(2 * k, dtype=np.int32)
    if peak_finder == "regionalmax":
        n = _fn("orc_find_peaks_regionalmax", prec)(_p(R), C.c_int(R.shape[0]), C.c_int(R.shape[1]),
                                                   C.c_int(k), _p(vals), _p(locs))
    else:
        n = _fn("orc_find_peaks_exclusion", prec)(_p(R), C.c_int(R.shape[0]), C.c_int(R.shape[1]),
                                                 C.c_int(k), C.c_int(exclusion_radius), _p(vals), _p(locs))
    return [(vals[m].item(), (int(locs[2 * m]), int(locs[2 * m + 1]))) for m in range(n)]


def gauss3(R, i, j, prec="f64"):
    dt = _dt(prec)
    R = _f(R, dt)
    out = np.zeros(2, dtype=dt)
    _fn("orc_gauss3", prec)(_p(R), C.c_int(R.shape[0]), C.c_int(R.shape[1]), C.c_int(i), C.c_int(j), _p(out))
    return float(out[0]), float(out[1])


def gauss9(R, i, j, prec="f64"):
    dt = _dt(prec)
    R = _f(R, dt)
    out = np.zeros(2, dtype=dt)
    _fn("orc_gauss9", prec)(_p(R), C.c_int(R.shape[0]), C.c_int(R.shape[1]), C.c_int(i), C.c_int(j), _p(out))
    return float(out[0]), float(out[1])


def moment(R, pr, pc, prec="f64"):
    dt = _dt(prec)
    R = _f(R, dt)
    fn = _fn("orc_moment", prec)
    fn.restype = C.c_double if prec == "f64" else C.c_float
    ct = C.c_double if prec == "f64" else C.c_float
    return float(fn(_p(R), C.c_int(R.shape[0]), C.c_int(R.shape[1]), ct(pr), ct(pc)))


def analyze_plane(R, pass_struct, prec="f64"):
    """analyze_plane! + alternatives: (du, dv, ratio, moment, found, alt_du, alt_dv)."""
    dt = _dt(prec)
    R = _f(R, dt)
    K = max(pass_struct.n_peaks, 2)
    vals = np.zeros(K, dtype=dt)
    locs = np.zeros(2 * K, dtype=np.int32)
    out = np.zeros(4, dtype=dt)
    na = max(pass_struct.n_peaks - 1, 1)
    au = np.full(na, np.nan, dtype=dt)
    av = np.full(na, np.nan, dtype=dt)
    found = _fn("orc_analyze_plane", prec)(_p(R), C.c_int(R.shape[0]), C.c_int(R.shape[1]),
                                          C.byref(pass_struct), _p(vals), _p(locs), _p(out), _p(au), _p(av))
    return (float(out[0]), float(out[1]), float(out[2]), float(out[3]), int(found), au, av)


def median(v, prec="f64"):
    dt = _dt(prec)
    a = np.array(v, dtype=dt)
    fn = _fn("orc_median", prec)
    fn.restype = C.c_double if prec == "f64" else C.c_float
    return float(fn(_p(a), C.c_int(a.size)))


def quantile(v, p, prec="f64"):
    dt = _dt(prec)
    a = np.array(v, dtype=dt)
    fn = _fn("orc_quantile", prec)
    fn.restype = C.c_double if prec == "f64" else C.c_float
    return float(fn(_p(a), C.c_int(a.size), C.c_double(p)))


def uod(u, v, threshold=2.0, neighborhood=1, epsilon=0.1, exclude=None, prec="f64"):
    dt = _dt(prec)
    u, v = _f(u, dt), _f(v, dt)
    flags = np.zeros(u.shape, dtype=np.uint8, order="F")
    ex = None if exclude is None else _f(exclude, np.uint8)
    _fn("orc_uod", prec)(_p(u), _p(v), C.c_int(u.shape[0]), C.c_int(u.shape[1]), C.c_double(threshold),
                         C.c_int(neighborhood), C.c_double(epsilon), _p(ex), _p(flags))
    return flags.astype(bool)


def replace_vectors(u, v, invalid, prec="f64"):
    dt = _dt(prec)
    u, v = _f(u, dt).copy(order="F"), _f(v, dt).copy(order="F")
    inv = _f(invalid, np.uint8)
    _fn("orc_replace", prec)(_p(u), _p(v), _p(inv), C.c_int(u.shape[0]), C.c_int(u.shape[1]))
    return u, v


def smooth_field(f, prec="f64"):
    dt = _dt(prec)
    f = _f(f, dt)
    out = np.zeros(f.shape, dtype=dt, order="F")
    _fn("orc_smooth", prec)(_p(f), C.c_int(f.shape[0]), C.c_int(f.shape[1]), _p(out))
    return out


def substitute(u, v, outliers, alt_u, alt_v, er=2, thr=2.0, mask=None, prec="f64"):
    dt = _dt(prec)
    u, v = _f(u, dt).copy(order="F"), _f(v, dt).copy(order="F")
    o = _f(outliers, np.uint8).copy(order="F")
    au, av = _f(alt_u, dt), _f(alt_v, dt)
    m = None if mask is None else _f(mask, np.uint8)
    n = _fn("orc_substitute", prec)(_p(u), _p(v), _p(o), _p(m), C.c_int(u.shape[0]), C.c_int(u.shape[1]),
                                    _p(au), _p(av), C.c_int(au.shape[2]), C.c_int(er), C.c_double(thr))
    return u, v, o.astype(bool), int(n)


def validate_and_replace(pass_struct, u, v, pr, cm, alt_u=None, alt_v=None, force_replace=False, mask=None,
                         prec="f64"):
    dt = _dt(prec)
    u, v = _f(u, dt).copy(order="F"), _f(v, dt).copy(order="F")
    pr, cm = _f(pr, dt), _f(cm, dt)
    au = None if alt_u is None else _f(alt_u, dt)
    av = None if alt_v is None else _f(alt_v, dt)
    m = None if mask is None else _f(mask, np.uint8)
    out = np.zeros(u.shape, dtype=np.uint8, order="F")
    _fn("orc_validate_and_replace", prec)(C.byref(pass_struct), C.c_int(u.shape[0]), C.c_int(u.shape[1]),
                                          C.c_int(1 if force_replace else 0), _p(u), _p(v), _p(pr), _p(cm),
                                          _p(au), _p(av), _p(m), _p(out))
    return u, v, out.astype(bool)


def prefilter(img, prec="f64"):
    dt = _dt(prec)
    img = _f(img, dt)
    nr, nc = img.shape
    coef = np.zeros((nr + 2, nc + 2), dtype=dt, order="F")
    _fn("orc_prefilter", prec)(_p(img), C.c_int(nr), C.c_int(nc), _p(coef))
    return coef


def cubic(coef, nr, nc, yy, xx, prec="f64"):
    dt = _dt(prec)
    coef = _f(coef, dt)
    fn = _fn("orc_cubic", prec)
    ct = C.c_double if prec == "f64" else C.c_float
    fn.restype = ct
    return float(fn(_p(coef), C.c_int(nr), C.c_int(nc), ct(yy), ct(xx)))


def deform(coefA, coefB, nr, nc, py, px, pu, pv, nthreads=8, prec="f64"):
    dt = _dt(prec)
    cA, cB = _f(coefA, dt), _f(coefB, dt)
    py, px = np.asarray(py, dtype=dt), np.asarray(px, dtype=dt)
    pu, pv = _f(pu, dt), _f(pv, dt)
    wA = np.zeros((nr, nc), dtype=dt, order="F")
    wB = np.zeros((nr, nc), dtype=dt, order="F")
    _fn("orc_deform", prec)(_p(cA), _p(cB), C.c_int(nr), C.c_int(nc), _p(py), C.c_int(py.size), _p(px),
                            C.c_int(px.size), _p(pu), _p(pv), _p(wA), _p(wB), C.c_int(nthreads))
    return wA, wB


def dewarp(img, rows, cols, mask=None, prec="f64"):
    """dewarp!(out, dw, img): the frame resampled at the dewarper's source
    coordinates (ny x nx, 1-based); masked nodes 0."""
    dt = _dt(prec)
    img = _f(img, dt)
    rows = np.asfortranarray(rows, dtype=np.float64)
    cols = np.asfortranarray(cols, dtype=np.float64)
    ny, nx = rows.shape
    out = np.zeros((ny, nx), dtype=dt, order="F")
    _fn("orc_dewarp", prec)(_p(img), C.c_int(img.shape[0]), C.c_int(img.shape[1]), _p(rows), _p(cols),
                            _p(_mask(mask)), C.c_int(ny), C.c_int(nx), _p(out))
    return out


def node_values(py, px, pu, pv, y, x, prec="f64"):
    dt = _dt(prec)
    py, px, y, x = (np.asarray(a, dtype=dt) for a in (py, px, y, x))
    pu, pv = _f(pu, dt), _f(pv, dt)
    u = np.zeros((y.size, x.size), dtype=dt, order="F")
    v = np.zeros((y.size, x.size), dtype=dt, order="F")
    _fn("orc_node_values", prec)(_p(py), C.c_int(py.size), _p(px), C.c_int(px.size), _p(pu), _p(pv), _p(y),
                                 C.c_int(y.size), _p(x), C.c_int(x.size), _p(u), _p(v))
    return u, v


def pass_grid(pass_struct, nr, nc, prec="f64"):
    dt = _dt(prec)
    ny, nx = C.c_int(), C.c_int()
    rc = _fn("orc_pass_grid", prec)(C.byref(pass_struct), C.c_int(nr), C.c_int(nc), C.byref(ny), C.byref(nx),
                                    None, None, None)
    if rc != 0:
        raise ValueError("search area exceeds image size")
    x = np.zeros(nx.value, dtype=dt)
    y = np.zeros(ny.value, dtype=dt)
    jobs = np.zeros((4, ny.value * nx.value), dtype=np.int32, order="F")
    _fn("orc_pass_grid", prec)(C.byref(pass_struct), C.c_int(nr), C.c_int(nc), C.byref(ny), C.byref(nx), _p(x),
                               _p(y), _p(jobs))
    return x, y, jobs


def process_windows(pass_struct, A, B, jobs, ny, nx, uncertainty=False, debug=False, planes=False, nthreads=8,
                    prec="f64", mask=None):
    """process_windows! over jobs (4 x n int32). Returns a dict of grids
    (residual u, v, peak_ratio, moment, alt_u, alt_v, unc_u, unc_v) and, with
    debug, per-job peak locs (2 x K x n), vals (K x n), found (n)."""
    dt = _dt(prec)
    A, B = _f(A, dt), _f(B, dt)
    jobs = _f(jobs, np.int32)
    nj = jobs.shape[1]
    K = max(pass_struct.n_peaks, 2)
    nalt = max(pass_struct.n_peaks - 1, 1)
    g = lambda: np.zeros((ny, nx), dtype=dt, order="F")
    res = dict(u=g(), v=g(), peak_ratio=g(), correlation_moment=g())
    res["alt_u"] = np.full((ny, nx, nalt), np.nan, dtype=dt, order="F")
    res["alt_v"] = np.full((ny, nx, nalt), np.nan, dtype=dt, order="F")
    uu = np.full((ny, nx), np.nan, dtype=dt, order="F") if uncertainty else None
    uv = np.full((ny, nx), np.nan, dtype=dt, order="F") if uncertainty else None
    locs = np.zeros((2, K, nj), dtype=np.int32, order="F") if debug else None
    vals = np.zeros((K, nj), dtype=dt, order="F") if debug else None
    found = np.zeros(nj, dtype=np.int32) if debug else None
    m1, m2 = _plane_shape(pass_struct)
    pl = np.zeros((m1, m2, nj), dtype=dt, order="F") if planes else None
    tie = np.ones((ny, nx), dtype=dt, order="F") if debug else None
    _fn("orc_process_windows", prec)(C.byref(pass_struct), _p(A), _p(B), C.c_int(A.shape[0]), C.c_int(A.shape[1]),
                                     _p(jobs), C.c_int(nj), C.c_int(ny), C.c_int(nx), _p(res["u"]), _p(res["v"]),
                                     _p(res["peak_ratio"]), _p(res["correlation_moment"]),
                                     _p(res["alt_u"]) if pass_struct.n_peaks > 1 else None,
                                     _p(res["alt_v"]) if pass_struct.n_peaks > 1 else None, _p(uu), _p(uv),
                                     _p(locs), _p(vals), _p(found), _p(pl), _p(_mask(mask)), C.c_int(nthreads),
                                     _p(tie))
    if uncertainty:
        res["unc_u"], res["unc_v"] = uu, uv
    if debug:
        res["locs"], res["vals"], res["found"] = locs, vals, found
        res["tie"] = tie
    if planes:
        res["planes"] = pl
    return res


def uncertainty_sweep(pass_struct, A, B, jobs, ny, nx, stats=False, nthreads=8, prec="f64", mask=None):
    dt = _dt(prec)
    A, B = _f(A, dt), _f(B, dt)
    jobs = _f(jobs, np.int32)
    nj = jobs.shape[1]
    uu = np.full((ny, nx), np.nan, dtype=dt, order="F")
    uv = np.full((ny, nx), np.nan, dtype=dt, order="F")
    st = np.zeros((44, 2, nj), dtype=np.float64, order="F") if stats else None
    _fn("orc_uncertainty_sweep", prec)(C.byref(pass_struct), _p(A), _p(B), C.c_int(A.shape[0]), _p(jobs),
                                       C.c_int(nj), C.c_int(ny), _p(uu), _p(uv), _p(st), _p(_mask(mask)),
                                       C.c_int(nthreads))
    return (uu, uv, st) if stats else (uu, uv)


def uq_finalize(stats44, prec="f64"):
    s = np.ascontiguousarray(stats44, dtype=np.float64)
    fn = _fn("orc_uq_finalize", prec)
    fn.restype = C.c_double if prec == "f64" else C.c_float
    return float(fn(_p(s)))


def uq_offsets():
    out = np.zeros(80, dtype=np.int32)
    rs = np.zeros(5, dtype=np.int32)
    n = lib().orc_uq_offset_table(_p(out), _p(rs))
    return out[: 2 * n].reshape(n, 2), rs


def field_change(u, v, pu, pv, prec="f64"):
    dt = _dt(prec)
    u, v, pu, pv = (np.ascontiguousarray(a, dtype=dt).ravel(order="F") for a in (u, v, pu, pv))
    fn = _fn("orc_field_change", prec)
    fn.restype = C.c_double if prec == "f64" else C.c_float
    return float(fn(_p(u), _p(v), _p(pu), _p(pv), None, C.c_int(u.size)))


def _mask(mask):
    return None if mask is None else np.asfortranarray(mask, dtype=np.uint8)


def grid_mask(pass_struct, nr, nc, mask, threshold=0.5):
    """pass_grid's grid-level mask (ny x nx bool) of an image-sized mask."""
    x, y, _ = pass_grid(pass_struct, nr, nc)
    gm = np.zeros((y.size, x.size), dtype=np.uint8, order="F")
    lib().orc_grid_mask_f64(C.byref(pass_struct), C.c_int(nr), C.c_int(nc), _p(_mask(mask)), C.c_double(threshold),
                            _p(gm))
    return gm.astype(bool)


def run_piv(A, B, passes, predictor_smoothing=True, nthreads=8, prec="f64", mask=None, mask_threshold=0.5,
            uq_pert=False, tie_rel=1e-5, max_ties=100000):
    """run_piv(imgA, imgB, passes; mask, mask_threshold): dict with the final
    PIVResult fields (``mask``: the final grid mask). With ``uq_pert`` and
    uncertainty on the final pass, ``uq_pert_u`` / ``uq_pert_v`` hold the
    final uncertainty recomputed on its input images rounded to float32 and
    scaled by 1 + 2^-21 h per pixel (h a fixed hash in [-1, 1]): sigma's own
    sensitivity to float32-level input error, and ``uq_stats`` the FP64
    Wieneke statistics behind each cell's sigma, shape (44, 2, ny, nx) (u then
    v). ``ties`` lists the windows of every pass and sweep whose primary peak
    is a near-tie (relative margin to a neighbour or to the secondary below
    ``tie_rel``) as rows (pass, sweep, y, x)."""
    dt = _dt(prec)
    A, B = _f(A, dt), _f(B, dt)
    nr, nc = A.shape
    arr = (type(passes[0]) * len(passes))(*passes)
    x, y, _ = pass_grid(passes[-1], nr, nc, prec)
    ny, nx = y.size, x.size
    g = lambda: np.zeros((ny, nx), dtype=dt, order="F")
    res = dict(x=np.zeros(nx, dtype=dt), y=np.zeros(ny, dtype=dt), u=g(), v=g(), peak_ratio=g(),
               correlation_moment=g(), uncertainty_u=g(), uncertainty_v=g(),
               outliers=np.zeros((ny, nx), dtype=np.uint8, order="F"),
               mask=np.zeros((ny, nx), dtype=np.uint8, order="F"))
    m = _mask(mask)
    st = np.zeros((ny, nx, 2), dtype=dt, order="F") if uq_pert else None
    ss = np.zeros((44, 2, ny, nx), dtype=np.float64, order="F") if uq_pert else None
    ties = np.zeros((max_ties, 4), dtype=np.float64)
    nties = C.c_int(0)
    rc = _fn("orc_run_piv", prec)(_p(A), _p(B), C.c_int(nr), C.c_int(nc), arr, C.c_int(len(passes)),
                                  C.c_int(1 if predictor_smoothing else 0), C.c_int(nthreads), _p(res["x"]),
                                  _p(res["y"]), _p(res["u"]), _p(res["v"]), _p(res["peak_ratio"]),
                                  _p(res["correlation_moment"]), _p(res["uncertainty_u"]),
                                  _p(res["uncertainty_v"]), _p(res["outliers"]), _p(m), C.c_double(mask_threshold),
                                  _p(res["mask"]), _p(st), _p(ss), C.c_double(tie_rel), _p(ties), C.c_int(max_ties),
                                  C.byref(nties))
    if rc != 0:
        raise ValueError("invalid pass schedule")
    if uq_pert:
        res["uq_pert_u"], res["uq_pert_v"] = st[:, :, 0], st[:, :, 1]
        res["uq_stats"] = ss
    res["ties"] = ties[:min(nties.value, max_ties)]
    res["ties_total"] = nties.value
    res["outliers"] = res["outliers"].astype(bool)
    res["mask"] = res["mask"].astype(bool)
    return res


def run_piv_ensemble(pairs, passes, predictor_smoothing=True, nthreads=8, prec="f64", planes=False, mask=None,
                     mask_threshold=0.5):
    """run_piv_ensemble(pairs, passes; mask, mask_threshold): dict with the
    final PIVResult fields (``mask``: the final grid mask); with ``planes``
    also the last pass's FP64 plane sums (m1 x m2 per unmasked window)."""
    dt = _dt(prec)
    As = [_f(A, dt) for A, _ in pairs]
    Bs = [_f(B, dt) for _, B in pairs]
    nr, nc = As[0].shape
    arr = (type(passes[0]) * len(passes))(*passes)
    x, y, jobs = pass_grid(passes[-1], nr, nc, prec)
    ny, nx = y.size, x.size
    g = lambda: np.zeros((ny, nx), dtype=dt, order="F")
    res = dict(x=np.zeros(nx, dtype=dt), y=np.zeros(ny, dtype=dt), u=g(), v=g(), peak_ratio=g(),
               correlation_moment=g(), uncertainty_u=g(), uncertainty_v=g(),
               outliers=np.zeros((ny, nx), dtype=np.uint8, order="F"),
               mask=np.zeros((ny, nx), dtype=np.uint8, order="F"))
    last = passes[-1]
    m1, m2 = (2 * s if last.padding else s for s in last.search_area_size)
    pl = np.zeros((m1, m2, ny * nx), dtype=np.float64, order="F") if planes else None
    pa = (C.c_void_p * len(As))(*[_p(A) for A in As])
    pb = (C.c_void_p * len(Bs))(*[_p(B) for B in Bs])
    rc = _fn("orc_run_piv_ensemble", prec)(pa, pb, C.c_int(len(As)), C.c_int(nr), C.c_int(nc), arr,
                                           C.c_int(len(passes)), C.c_int(1 if predictor_smoothing else 0),
                                           C.c_int(nthreads), _p(res["x"]), _p(res["y"]), _p(res["u"]),
                                           _p(res["v"]), _p(res["peak_ratio"]), _p(res["correlation_moment"]),
                                           _p(res["uncertainty_u"]), _p(res["uncertainty_v"]),
                                           _p(res["outliers"]), _p(pl), _p(_mask(mask)),
                                           C.c_double(mask_threshold), _p(res["mask"]))
    if rc != 0:
        raise ValueError("invalid pass schedule or empty pair list")
    res["outliers"] = res["outliers"].astype(bool)
    res["mask"] = res["mask"].astype(bool)
    if planes:
        res["planes"] = pl[:, :, :int((~res["mask"]).sum())]
    return res


def grad_energy(pass_struct, A, B, py, px, pu, pv, gradient=True, interp="cubic", prec="f64"):
    """Windowed correlation energy of the node field (pu, pv on knots py, px)
    and, with ``gradient``, its gradient: (E, gu, gv)."""
    dt = _dt(prec)
    A, B = _f(A, dt), _f(B, dt)
    py = np.ascontiguousarray(py, dtype=dt)
    px = np.ascontiguousarray(px, dtype=dt)
    pu, pv = _f(pu, dt), _f(pv, dt)
    gu = np.zeros_like(pu) if gradient else None
    gv = np.zeros_like(pv) if gradient else None
    fn = _fn("orc_grad_energy", prec)
    fn.restype = C.c_double
    E = fn(C.byref(pass_struct), _p(A), _p(B), C.c_int(A.shape[0]), C.c_int(A.shape[1]),
           C.c_int(1 if interp == "cubic" else 0), _p(py), C.c_int(py.size),
           _p(px), C.c_int(px.size), _p(pu), _p(pv), _p(gu), _p(gv))
    return (E, gu, gv) if gradient else E


def default_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))
