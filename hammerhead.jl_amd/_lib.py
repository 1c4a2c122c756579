"""ctypes binding of libhammerhead_mi355x.so (include/hammerhead_mi355x.h).

There is no CPU fallback: if the library is missing, or no HIP device is
present, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

from .params import ArgumentError, DimensionMismatch, hh_pass

PKG_DIR = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libhammerhead_mi355x.so"

HH_OK, HH_EINVAL, HH_EDIM, HH_EUNSUPPORTED, HH_EDEVICE, HH_ENOMEM, HH_ESTATE = range(7)
HH_SRC_RAW, HH_SRC_WARP = 0, 1
HH_FORM_FFT32, HH_FORM_MFMA16 = 0, 1
UQ_NSTATS = 44

# every entry point of include/hammerhead_mi355x.h
EXPORTS = (
    "hh_create", "hh_destroy", "hh_last_error", "hh_version", "hh_set_form", "hh_check_passes", "hh_grid_size",
    "hh_set_pair", "hh_set_pair_device", "hh_prefilter", "hh_set_coefs", "hh_deform", "hh_get_images", "hh_get_coefs",
    "hh_process_windows", "hh_correlation_planes", "hh_uncertainty_sweep", "hh_validate_and_replace",
    "hh_validate_and_replace_noalloc",
    "hh_smooth_field", "hh_run_piv", "hh_fetch_result", "hh_ensemble_begin", "hh_ensemble_add",
    "hh_ensemble_analyze", "hh_ensemble_planes", "hh_run_piv_ensemble", "hh_energy", "hh_grad_energy",
    "hh_refine_nodes", "hh_set_energy_sampling", "hh_stream", "hh_synchronize", "hh_set_timing", "hh_set_timing_only", "hh_kernel_time",
    "hh_set_pairs", "hh_run_piv_batch", "hh_fetch_pair", "hh_pair_grid", "hh_set_graphs", "hh_graph_replays",
    "hh_set_autotune", "hh_tune_report", "hh_tune_set", "hh_tune_set_family", "hh_set_mask", "hh_set_dewarp_map", "hh_dewarp",
    "hh_dewarp_pair",
)


class DeviceError(RuntimeError):
    """A HIP runtime failure or call-order error reported by the library."""


class hh_field(C.Structure):
    _fields_ = [
        ("ny", C.c_int32), ("nx", C.c_int32),
        ("x", C.c_void_p), ("y", C.c_void_p),
        ("u", C.c_void_p), ("v", C.c_void_p), ("peak_ratio", C.c_void_p), ("correlation_moment", C.c_void_p),
        ("uncertainty_u", C.c_void_p), ("uncertainty_v", C.c_void_p),
        ("outliers", C.c_void_p), ("mask", C.c_void_p),
    ]


_lib = None


def load():
    """Load the shared library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: build it with hammerhead.jl_amd/csrc/build.sh "
                          "(or __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(str(LIB_PATH))
    vp, i32, pp = C.c_void_p, C.c_int, C.POINTER(hh_pass)
    sig = {
        "hh_create": (i32, [C.POINTER(vp), i32]),
        "hh_destroy": (i32, [vp]),
        "hh_last_error": (C.c_char_p, [vp]),
        "hh_version": (C.c_char_p, []),
        "hh_set_form": (i32, [vp, i32]),
        "hh_check_passes": (i32, [vp, pp, i32]),
        "hh_grid_size": (i32, [pp, i32, i32, C.POINTER(i32), C.POINTER(i32)]),
        "hh_set_pair": (i32, [vp, vp, vp, i32, i32]),
        "hh_set_pair_device": (i32, [vp, vp, vp, i32, i32]),
        "hh_prefilter": (i32, [vp]),
        "hh_set_coefs": (i32, [vp, vp, vp, i32, i32]),
        "hh_deform": (i32, [vp, vp, i32, vp, i32, vp, vp]),
        "hh_get_images": (i32, [vp, i32, vp, vp]),
        "hh_get_coefs": (i32, [vp, vp, vp]),
        "hh_process_windows": (i32, [vp, pp, i32, vp, i32, i32, i32] + [vp] * 11),
        "hh_correlation_planes": (i32, [vp, pp, i32, vp, i32, vp]),
        "hh_uncertainty_sweep": (i32, [vp, pp, i32, vp, i32, i32, i32, vp, vp, vp]),
        "hh_validate_and_replace": (i32, [vp, pp, i32, i32, i32] + [vp] * 8),
        "hh_validate_and_replace_noalloc": (i32, [vp, pp, i32, i32, i32] + [vp] * 8),
        "hh_smooth_field": (i32, [vp, i32, i32, vp, vp]),
        "hh_run_piv": (i32, [vp, pp, i32, i32, C.POINTER(hh_field)]),
        "hh_fetch_result": (i32, [vp, C.POINTER(hh_field)]),
        "hh_ensemble_begin": (i32, [vp, pp, vp, i32, i32]),
        "hh_ensemble_add": (i32, [vp, i32]),
        "hh_ensemble_analyze": (i32, [vp, i32, i32] + [vp] * 8),
        "hh_ensemble_planes": (i32, [vp, vp]),
        "hh_run_piv_ensemble": (i32, [vp, vp, vp, i32, i32, i32, pp, i32, i32, C.POINTER(hh_field)]),
        "hh_energy": (i32, [vp, pp, vp, i32, vp, i32, vp, vp, C.POINTER(C.c_double)]),
        "hh_grad_energy": (i32, [vp, pp, vp, i32, vp, i32, vp, vp, C.POINTER(C.c_double), vp, vp]),
        "hh_refine_nodes": (i32, [vp, pp, vp, i32, vp, i32, vp, vp, i32, vp]),
        "hh_set_energy_sampling": (i32, [vp, i32]),
        "hh_stream": (vp, [vp]),
        "hh_synchronize": (i32, [vp]),
        "hh_set_timing": (i32, [vp, i32]),
        "hh_set_timing_only": (i32, [vp, C.c_char_p, i32]),
        "hh_kernel_time": (i32, [vp, C.c_char_p, i32, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
        "hh_set_pairs": (i32, [vp, vp, vp, vp, vp, i32]),
        "hh_run_piv_batch": (i32, [vp, pp, i32, i32, vp]),
        "hh_fetch_pair": (i32, [vp, i32, C.POINTER(hh_field)]),
        "hh_pair_grid": (i32, [vp, i32, C.POINTER(i32), C.POINTER(i32)]),
        "hh_set_graphs": (i32, [vp, i32]),
        "hh_graph_replays": (i32, [vp, C.POINTER(C.c_int64)]),
        "hh_set_autotune": (i32, [vp, i32]),
        "hh_tune_report": (i32, [vp, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
        "hh_tune_set": (i32, [vp, C.c_char_p, i32, i32]),
        "hh_tune_set_family": (i32, [vp, C.c_char_p, i32]),
        "hh_set_mask": (i32, [vp, vp, i32, i32, C.c_double]),
        "hh_set_dewarp_map": (i32, [vp, vp, vp, vp, i32, i32, i32, i32]),
        "hh_dewarp": (i32, [vp, vp, i32, i32, vp]),
        "hh_dewarp_pair": (i32, [vp, vp, vp, i32, i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(ctx_handle, rc, what):
    if rc == HH_OK:
        return
    msg = load().hh_last_error(ctx_handle)
    msg = msg.decode() if msg else ""
    text = f"{what}: {msg}" if msg else what
    if rc in (HH_EINVAL, HH_EUNSUPPORTED):
        raise ArgumentError(text)
    if rc == HH_EDIM:
        raise DimensionMismatch(text)
    raise DeviceError(f"{text} (status {rc})")


def _f32(a):
    return np.asfortranarray(a, dtype=np.float32)


def pass_array(passes):
    structs = [p.to_struct() if hasattr(p, "to_struct") else p for p in passes]
    return (hh_pass * len(structs))(*structs), len(structs)


class Context:
    """An hh_ctx: device buffers, engine pool and one HIP stream — the
    PIVWorkspace of a device backend (src/pipeline.jl:21-47)."""

    def __init__(self, device: int = 0):
        lib = load()
        h = C.c_void_p()
        rc = lib.hh_create(C.byref(h), device)
        if rc != HH_OK:
            _check(None, rc, "hh_create")
        self.h = h
        self.nr = self.nc = 0

    def close(self):
        if getattr(self, "h", None):
            load().hh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- configuration -------------------------------------------------
    def check_passes(self, passes):
        arr, n = pass_array(passes)
        _check(self.h, load().hh_check_passes(self.h, arr, n), "check_passes")

    def set_form(self, form: int):
        _check(self.h, load().hh_set_form(self.h, form), "set_form")

    # ---- image pair ----------------------------------------------------
    def set_pair(self, A, B):
        A, B = _f32(A), _f32(B)
        if A.shape != B.shape:
            raise DimensionMismatch(f"images must have the same size, got {A.shape} and {B.shape}")
        self.nr, self.nc = A.shape
        _check(self.h, load().hh_set_pair(self.h, _ptr(A), _ptr(B), A.shape[0], A.shape[1]), "set_pair")

    def set_pair_device(self, dA_ptr: int, dB_ptr: int, nr: int, nc: int):
        self.nr, self.nc = nr, nc
        _check(self.h, load().hh_set_pair_device(self.h, C.c_void_p(dA_ptr), C.c_void_p(dB_ptr), nr, nc),
               "set_pair_device")

    def set_mask(self, mask=None, mask_threshold=0.5):
        """The mask / mask_threshold of every following run_piv and
        run_piv_batch (None removes it); an identical mask keeps the plan."""
        if mask is None:
            _check(self.h, load().hh_set_mask(self.h, None, 0, 0, float(mask_threshold)), "set_mask")
            return
        m = np.asfortranarray(np.asarray(mask, dtype=bool), dtype=np.uint8)
        _check(self.h, load().hh_set_mask(self.h, _ptr(m), m.shape[0], m.shape[1], float(mask_threshold)),
               "set_mask")

    # ---- dewarping ---------------------------------------------------------
    def set_dewarp_map(self, rows, cols, mask=None, image_size=None):
        """The precomputed fields of an ImageDewarper: rows, cols (ny x nx,
        1-based source coordinates), out-of-view mask, camera frame size."""
        rows = np.asfortranarray(rows, dtype=np.float64)
        cols = np.asfortranarray(cols, dtype=np.float64)
        if rows.shape != cols.shape:
            raise DimensionMismatch(f"rows {rows.shape} and cols {cols.shape} differ")
        m = None if mask is None else np.asfortranarray(np.asarray(mask, dtype=bool), dtype=np.uint8)
        nr, nc = image_size
        self.dw_shape = rows.shape
        _check(self.h, load().hh_set_dewarp_map(self.h, _ptr(rows), _ptr(cols), _ptr(m), rows.shape[0],
                                                 rows.shape[1], nr, nc), "set_dewarp_map")

    def dewarp(self, img):
        img = _f32(img)
        out = np.zeros(self.dw_shape, dtype=np.float32, order="F")
        _check(self.h, load().hh_dewarp(self.h, _ptr(img), img.shape[0], img.shape[1], _ptr(out)), "dewarp")
        return out

    def dewarp_pair(self, A, B):
        """Dewarp both frames; the result becomes the current (device) pair."""
        A, B = _f32(A), _f32(B)
        if A.shape != B.shape:
            raise DimensionMismatch(f"images must have the same size, got {A.shape} and {B.shape}")
        _check(self.h, load().hh_dewarp_pair(self.h, _ptr(A), _ptr(B), A.shape[0], A.shape[1]), "dewarp_pair")
        self.nr, self.nc = self.dw_shape

    def prefilter(self):
        _check(self.h, load().hh_prefilter(self.h), "prefilter")

    def set_coefs(self, cA, cB):
        """Stage prefiltered coefficient arrays ((nr+2) x (nc+2) each)."""
        cA, cB = _f32(cA), _f32(cB)
        if cA.shape != cB.shape:
            raise DimensionMismatch(f"coefficient arrays differ: {cA.shape} and {cB.shape}")
        self.nr, self.nc = cA.shape[0] - 2, cA.shape[1] - 2
        _check(self.h, load().hh_set_coefs(self.h, _ptr(cA), _ptr(cB), self.nr, self.nc), "set_coefs")

    def coefs(self):
        cA = np.zeros((self.nr + 2, self.nc + 2), dtype=np.float32, order="F")
        cB = np.zeros_like(cA)
        _check(self.h, load().hh_get_coefs(self.h, _ptr(cA), _ptr(cB)), "get_coefs")
        return cA, cB

    def deform(self, py, px, pu, pv):
        py = np.ascontiguousarray(py, dtype=np.float32)
        px = np.ascontiguousarray(px, dtype=np.float32)
        pu, pv = _f32(pu), _f32(pv)
        _check(self.h, load().hh_deform(self.h, _ptr(py), py.size, _ptr(px), px.size, _ptr(pu), _ptr(pv)), "deform")

    def images(self, src=HH_SRC_WARP):
        A = np.zeros((self.nr, self.nc), dtype=np.float32, order="F")
        B = np.zeros_like(A)
        _check(self.h, load().hh_get_images(self.h, src, _ptr(A), _ptr(B)), "get_images")
        return A, B

    # ---- hooks -----------------------------------------------------------
    def process_windows(self, p, jobs, ny, nx, src=HH_SRC_RAW, uncertainty=False, debug=False, out=None):
        """process_windows!: residual grids (NaN outside jobs unless `out` given)."""
        st = p.to_struct() if hasattr(p, "to_struct") else p
        jobs = np.asfortranarray(jobs, dtype=np.int32)
        nj = jobs.shape[1] if jobs.ndim == 2 else 0
        K = max(st.n_peaks, 2)
        nalt = max(st.n_peaks - 1, 1)
        if out is None:
            g = lambda: np.full((ny, nx), np.nan, dtype=np.float32, order="F")
            out = dict(u=g(), v=g(), peak_ratio=g(), correlation_moment=g(),
                       alt_u=np.full((ny, nx, nalt), np.nan, dtype=np.float32, order="F"),
                       alt_v=np.full((ny, nx, nalt), np.nan, dtype=np.float32, order="F"))
            if uncertainty:
                out["unc_u"], out["unc_v"] = g(), g()
        if debug:
            out["locs"] = np.zeros((2, K, nj), dtype=np.int32, order="F")
            out["vals"] = np.zeros((K, nj), dtype=np.float32, order="F")
            out["found"] = np.zeros(nj, dtype=np.int32)
        has_alt = st.n_peaks > 1
        rc = load().hh_process_windows(
            self.h, C.byref(st), src, _ptr(jobs), nj, ny, nx, _ptr(out["u"]), _ptr(out["v"]),
            _ptr(out["peak_ratio"]), _ptr(out["correlation_moment"]),
            _ptr(out["alt_u"]) if has_alt else None, _ptr(out["alt_v"]) if has_alt else None,
            _ptr(out.get("unc_u")) if uncertainty else None, _ptr(out.get("unc_v")) if uncertainty else None,
            _ptr(out.get("locs")) if debug else None, _ptr(out.get("vals")) if debug else None,
            _ptr(out.get("found")) if debug else None)
        _check(self.h, rc, "process_windows")
        return out

    def correlation_planes(self, p, jobs, src=HH_SRC_RAW):
        st = p.to_struct() if hasattr(p, "to_struct") else p
        jobs = np.asfortranarray(jobs, dtype=np.int32)
        sz = st.search_area_size[0]
        M = 2 * sz if st.padding else sz
        planes = np.zeros((M, M, jobs.shape[1]), dtype=np.float32, order="F")
        _check(self.h, load().hh_correlation_planes(self.h, C.byref(st), src, _ptr(jobs), jobs.shape[1],
                                                     _ptr(planes)), "correlation_planes")
        return planes

    def uncertainty_sweep(self, p, jobs, ny, nx, src=HH_SRC_WARP, stats=False):
        st = p.to_struct() if hasattr(p, "to_struct") else p
        jobs = np.asfortranarray(jobs, dtype=np.int32)
        uu = np.full((ny, nx), np.nan, dtype=np.float32, order="F")
        uv = np.full((ny, nx), np.nan, dtype=np.float32, order="F")
        s = np.zeros((UQ_NSTATS, 2, jobs.shape[1]), dtype=np.float64, order="F") if stats else None
        _check(self.h, load().hh_uncertainty_sweep(self.h, C.byref(st), src, _ptr(jobs), jobs.shape[1], ny, nx,
                                                    _ptr(uu), _ptr(uv), _ptr(s)), "uncertainty_sweep")
        return (uu, uv, s) if stats else (uu, uv)

    def validate_and_replace(self, p, u, v, peak_ratio, correlation_moment, alt_u=None, alt_v=None,
                             force_replace=False, mask=None):
        st = p.to_struct() if hasattr(p, "to_struct") else p
        u, v = _f32(u).copy(order="F"), _f32(v).copy(order="F")
        pr, cm = _f32(peak_ratio), _f32(correlation_moment)
        au = None if alt_u is None else _f32(alt_u)
        av = None if alt_v is None else _f32(alt_v)
        m = None if mask is None else np.asfortranarray(mask, dtype=np.uint8)
        out = np.zeros(u.shape, dtype=np.uint8, order="F")
        _check(self.h, load().hh_validate_and_replace(self.h, C.byref(st), u.shape[0], u.shape[1],
                                                       1 if force_replace else 0, _ptr(u), _ptr(v), _ptr(pr),
                                                       _ptr(cm), _ptr(au), _ptr(av), _ptr(m), _ptr(out)),
               "validate_and_replace")
        return u, v, out.astype(bool)

    def validate_and_replace_noalloc(self, p, ny, nx, d_u, d_v, d_pr, d_cm, d_alt_u=None, d_alt_v=None,
                                     d_mask=None, d_outliers=None, force_replace=False):
        """hh_validate_and_replace_noalloc on device pointers (ints); enqueued
        on the context stream, returns without waiting."""
        st = p.to_struct() if hasattr(p, "to_struct") else p
        vp = lambda x: None if x is None else C.c_void_p(x)
        _check(self.h, load().hh_validate_and_replace_noalloc(self.h, C.byref(st), ny, nx, 1 if force_replace else 0,
                                                               vp(d_u), vp(d_v), vp(d_pr), vp(d_cm), vp(d_alt_u),
                                                               vp(d_alt_v), vp(d_mask), vp(d_outliers)),
               "validate_and_replace_noalloc")

    def smooth_field(self, f):
        f = _f32(f)
        out = np.zeros_like(f, order="F")
        _check(self.h, load().hh_smooth_field(self.h, f.shape[0], f.shape[1], _ptr(f), _ptr(out)), "smooth_field")
        return out

    # ---- whole schedule ----------------------------------------------------
    def run_piv(self, passes, predictor_smoothing=True, fetch=True):
        arr, n = pass_array(passes)
        if not fetch:
            _check(self.h, load().hh_run_piv(self.h, arr, n, 1 if predictor_smoothing else 0, None), "run_piv")
            return None
        ny, nx = grid_size(passes[-1], self.nr, self.nc)
        res = dict(x=np.zeros(nx, np.float32), y=np.zeros(ny, np.float32))
        for k in ("u", "v", "peak_ratio", "correlation_moment", "uncertainty_u", "uncertainty_v"):
            res[k] = np.zeros((ny, nx), dtype=np.float32, order="F")
        res["outliers"] = np.zeros((ny, nx), dtype=np.uint8, order="F")
        res["mask"] = np.zeros((ny, nx), dtype=np.uint8, order="F")
        f = hh_field(ny, nx, *(_ptr(res[k]) for k in ("x", "y", "u", "v", "peak_ratio", "correlation_moment",
                                                       "uncertainty_u", "uncertainty_v", "outliers", "mask")))
        _check(self.h, load().hh_run_piv(self.h, arr, n, 1 if predictor_smoothing else 0, C.byref(f)), "run_piv")
        res["outliers"] = res["outliers"].astype(bool)
        res["mask"] = res["mask"].astype(bool)
        return res

    def _result_arrays(self, ny, nx):
        res = dict(x=np.zeros(nx, np.float32), y=np.zeros(ny, np.float32))
        for k in ("u", "v", "peak_ratio", "correlation_moment", "uncertainty_u", "uncertainty_v"):
            res[k] = np.zeros((ny, nx), dtype=np.float32, order="F")
        res["outliers"] = np.zeros((ny, nx), dtype=np.uint8, order="F")
        res["mask"] = np.zeros((ny, nx), dtype=np.uint8, order="F")
        f = hh_field(ny, nx, *(_ptr(res[k]) for k in ("x", "y", "u", "v", "peak_ratio", "correlation_moment",
                                                       "uncertainty_u", "uncertainty_v", "outliers", "mask")))
        return res, f

    def fetch_result(self, ny, nx):
        res, f = self._result_arrays(ny, nx)
        _check(self.h, load().hh_fetch_result(self.h, C.byref(f)), "fetch_result")
        res["outliers"] = res["outliers"].astype(bool)
        res["mask"] = res["mask"].astype(bool)
        return res

    # ---- batches of pairs of uneven sizes --------------------------------------
    def set_pairs(self, pairs):
        """The frames of run_piv_sequence: a list of (A, B) of any sizes,
        copied into one device-resident atlas."""
        As = [_f32(A) for A, _ in pairs]
        Bs = [_f32(B) for _, B in pairs]
        if not As:
            raise ArgumentError("pairs must not be empty")
        for A, B in zip(As, Bs):
            if A.shape != B.shape:
                raise DimensionMismatch(f"images must have the same size, got {A.shape} and {B.shape}")
        nr = np.array([A.shape[0] for A in As], dtype=np.int32)
        nc = np.array([A.shape[1] for A in As], dtype=np.int32)
        pa = (C.c_void_p * len(As))(*[A.ctypes.data for A in As])
        pb = (C.c_void_p * len(Bs))(*[B.ctypes.data for B in Bs])
        _check(self.h, load().hh_set_pairs(self.h, pa, pb, _ptr(nr), _ptr(nc), len(As)), "set_pairs")
        self.batch_shapes = list(zip(nr.tolist(), nc.tolist()))
        self.nr = self.nc = 0

    def run_piv_batch(self, passes, predictor_smoothing=True, fetch=True):
        """run_piv of every pair of the batch; a list of result dicts."""
        arr, n = pass_array(passes)
        _check(self.h, load().hh_run_piv_batch(self.h, arr, n, 1 if predictor_smoothing else 0, None),
               "run_piv_batch")
        if not fetch:
            return None
        return [self.fetch_pair(q) for q in range(len(self.batch_shapes))]

    def pair_grid(self, q):
        ny, nx = C.c_int(), C.c_int()
        _check(self.h, load().hh_pair_grid(self.h, q, C.byref(ny), C.byref(nx)), "pair_grid")
        return ny.value, nx.value

    def fetch_pair(self, q):
        ny, nx = self.pair_grid(q)
        res, f = self._result_arrays(ny, nx)
        _check(self.h, load().hh_fetch_pair(self.h, q, C.byref(f)), "fetch_pair")
        res["outliers"] = res["outliers"].astype(bool)
        res["mask"] = res["mask"].astype(bool)
        return res

    def set_graphs(self, enable: bool):
        _check(self.h, load().hh_set_graphs(self.h, 1 if enable else 0), "set_graphs")

    def set_autotune(self, enable, level: int = None):
        """enable: False (off) / True (launch geometry, level 1); level 2 also
        chooses the kernel family per FFT side."""
        lv = (0 if not enable else 1) if level is None else int(level)
        _check(self.h, load().hh_set_autotune(self.h, lv), "set_autotune")

    def tune_report(self) -> dict:
        import json

        need = C.c_size_t()
        _check(self.h, load().hh_tune_report(self.h, None, 0, C.byref(need)), "tune_report")
        buf = C.create_string_buffer(need.value)
        _check(self.h, load().hh_tune_report(self.h, buf, need.value, None), "tune_report")
        return json.loads(buf.value.decode())

    def tune_load(self, report: dict):
        """Preload the choices of a tune_report() (no candidate timing)."""
        for key, e in report.items():
            ch = e["choice"]
            _check(self.h, load().hh_tune_set(self.h, key.encode(), int(ch["waves_per_simd"]),
                                              int(ch["linear_order"])), "tune_set")
            _check(self.h, load().hh_tune_set_family(self.h, key.encode(), int(ch.get("family", 0))),
                   "tune_set_family")

    def graph_replays(self) -> int:
        n = C.c_int64()
        _check(self.h, load().hh_graph_replays(self.h, C.byref(n)), "graph_replays")
        return n.value

    # ---- ensemble ------------------------------------------------------------
    def ensemble_begin(self, p, jobs, uncertainty=False):
        st = p.to_struct() if hasattr(p, "to_struct") else p
        jobs = np.asfortranarray(jobs, dtype=np.int32)
        self._ens = (st, jobs.shape[1], bool(uncertainty))
        _check(self.h, load().hh_ensemble_begin(self.h, C.byref(st), _ptr(jobs), jobs.shape[1],
                                                 1 if uncertainty else 0), "ensemble_begin")

    def ensemble_add(self, src=HH_SRC_RAW):
        _check(self.h, load().hh_ensemble_add(self.h, src), "ensemble_add")

    def ensemble_analyze(self, ny, nx, u=None, v=None):
        """ensemble_analyze!: u, v = the shared predictor at the nodes (zeros
        when None); returns the total displacement and the quality grids."""
        st, nj, unc = self._ens
        g = lambda: np.zeros((ny, nx), dtype=np.float32, order="F")
        res = dict(u=g() if u is None else _f32(u).copy(order="F"), v=g() if v is None else _f32(v).copy(order="F"),
                   peak_ratio=g(), correlation_moment=g())
        nalt = max(st.n_peaks - 1, 1)
        res["alt_u"] = np.full((ny, nx, nalt), np.nan, dtype=np.float32, order="F")
        res["alt_v"] = np.full((ny, nx, nalt), np.nan, dtype=np.float32, order="F")
        res["uncertainty_u"] = np.full((ny, nx), np.nan, dtype=np.float32, order="F")
        res["uncertainty_v"] = np.full((ny, nx), np.nan, dtype=np.float32, order="F")
        has_alt = st.n_peaks > 1
        _check(self.h, load().hh_ensemble_analyze(
            self.h, ny, nx, _ptr(res["u"]), _ptr(res["v"]), _ptr(res["peak_ratio"]), _ptr(res["correlation_moment"]),
            _ptr(res["alt_u"]) if has_alt else None, _ptr(res["alt_v"]) if has_alt else None,
            _ptr(res["uncertainty_u"]), _ptr(res["uncertainty_v"])), "ensemble_analyze")
        return res

    def ensemble_planes(self):
        st, nj, _ = self._ens
        M = 2 * st.search_area_size[0] if st.padding else st.search_area_size[0]
        pl = np.zeros((M, M, nj), dtype=np.float64, order="F")
        _check(self.h, load().hh_ensemble_planes(self.h, _ptr(pl)), "ensemble_planes")
        return pl

    def run_piv_ensemble(self, pairs, passes, predictor_smoothing=True, fetch=True):
        """run_piv_ensemble(pairs, passes) with every pass on the device."""
        As = [_f32(A) for A, _ in pairs]
        Bs = [_f32(B) for _, B in pairs]
        if not As:
            raise ArgumentError("pairs must not be empty")
        shape = As[0].shape
        for A, B in zip(As, Bs):
            if A.shape != shape or B.shape != shape:
                raise DimensionMismatch(f"all pairs must share the image size {shape}")
        pa = (C.c_void_p * len(As))(*[A.ctypes.data for A in As])
        pb = (C.c_void_p * len(Bs))(*[B.ctypes.data for B in Bs])
        arr, n = pass_array(passes)
        self.nr, self.nc = shape
        if not fetch:
            _check(self.h, load().hh_run_piv_ensemble(self.h, pa, pb, len(As), shape[0], shape[1], arr, n,
                                                       1 if predictor_smoothing else 0, None), "run_piv_ensemble")
            return None
        ny, nx = grid_size(passes[-1], *shape)
        res, f = self._result_arrays(ny, nx)
        _check(self.h, load().hh_run_piv_ensemble(self.h, pa, pb, len(As), shape[0], shape[1], arr, n,
                                                   1 if predictor_smoothing else 0, C.byref(f)), "run_piv_ensemble")
        res["outliers"] = res["outliers"].astype(bool)
        res["mask"] = res["mask"].astype(bool)
        return res

    def run_piv_ensemble_resident(self, npairs, shape, passes, predictor_smoothing=True, fetch=True):
        """run_piv_ensemble again on the pairs the previous call uploaded."""
        arr, n = pass_array(passes)
        nr, nc = shape
        self.nr, self.nc = nr, nc
        if not fetch:
            _check(self.h, load().hh_run_piv_ensemble(self.h, None, None, npairs, nr, nc, arr, n,
                                                       1 if predictor_smoothing else 0, None), "run_piv_ensemble")
            return None
        ny, nx = grid_size(passes[-1], nr, nc)
        res, f = self._result_arrays(ny, nx)
        _check(self.h, load().hh_run_piv_ensemble(self.h, None, None, npairs, nr, nc, arr, n,
                                                   1 if predictor_smoothing else 0, C.byref(f)), "run_piv_ensemble")
        res["outliers"] = res["outliers"].astype(bool)
        res["mask"] = res["mask"].astype(bool)
        return res

    # ---- correlation energy and its adjoint ------------------------------------
    @staticmethod
    def _nodes(py, px, pu, pv):
        py = np.ascontiguousarray(py, dtype=np.float32)
        px = np.ascontiguousarray(px, dtype=np.float32)
        pu, pv = _f32(pu).copy(order="F"), _f32(pv).copy(order="F")
        if pu.shape != (py.size, px.size) or pv.shape != pu.shape:
            raise DimensionMismatch(f"node field must be {(py.size, px.size)}, got {pu.shape} and {pv.shape}")
        return py, px, pu, pv

    def set_energy_sampling(self, sampling: str):
        """Image sampling of the correlation energy: "cubic" (the
        deformation path's interpolant) or "bilinear"."""
        code = {"cubic": 0, "bilinear": 1}[sampling]
        _check(self.h, load().hh_set_energy_sampling(self.h, code), "set_energy_sampling")

    def energy(self, p, py, px, pu, pv):
        st = p.to_struct() if hasattr(p, "to_struct") else p
        py, px, pu, pv = self._nodes(py, px, pu, pv)
        E = C.c_double()
        _check(self.h, load().hh_energy(self.h, C.byref(st), _ptr(py), py.size, _ptr(px), px.size, _ptr(pu), _ptr(pv),
                                         C.byref(E)), "energy")
        return E.value

    def grad_energy(self, p, py, px, pu, pv):
        """(E, dE/dpu, dE/dpv) by the hand-written adjoint."""
        st = p.to_struct() if hasattr(p, "to_struct") else p
        py, px, pu, pv = self._nodes(py, px, pu, pv)
        E = C.c_double()
        gu, gv = np.zeros_like(pu), np.zeros_like(pv)
        _check(self.h, load().hh_grad_energy(self.h, C.byref(st), _ptr(py), py.size, _ptr(px), px.size, _ptr(pu),
                                              _ptr(pv), C.byref(E), _ptr(gu), _ptr(gv)), "grad_energy")
        return E.value, gu, gv

    def refine_nodes(self, p, py, px, pu, pv, niter=20):
        """niter gradient steps on the node field: (pu, pv, energies)."""
        st = p.to_struct() if hasattr(p, "to_struct") else p
        py, px, pu, pv = self._nodes(py, px, pu, pv)
        en = np.zeros(niter + 1, dtype=np.float64)
        _check(self.h, load().hh_refine_nodes(self.h, C.byref(st), _ptr(py), py.size, _ptr(px), px.size, _ptr(pu),
                                               _ptr(pv), niter, _ptr(en)), "refine_nodes")
        return pu, pv, en

    # ---- stream and timing -------------------------------------------------
    def stream(self) -> int:
        return load().hh_stream(self.h) or 0

    def synchronize(self):
        _check(self.h, load().hh_synchronize(self.h), "synchronize")

    def set_timing(self, enable: bool, only=None):
        """Per-kernel HIP-event timing; only=(family, window) times that
        kernel family alone."""
        if enable and only is not None:
            _check(self.h, load().hh_set_timing_only(self.h, only[0].encode(), int(only[1])), "set_timing")
        else:
            _check(self.h, load().hh_set_timing(self.h, 1 if enable else 0), "set_timing")

    def kernel_time(self, family: str, window: int):
        ms, n = C.c_double(), C.c_int64()
        _check(self.h, load().hh_kernel_time(self.h, family.encode(), window, C.byref(ms), C.byref(n)),
               "kernel_time")
        return ms.value, n.value


def grid_size(p, nr, nc):
    st = p.to_struct() if hasattr(p, "to_struct") else p
    ny, nx = C.c_int(), C.c_int()
    rc = load().hh_grid_size(C.byref(st), nr, nc, C.byref(ny), C.byref(nx))
    if rc != HH_OK:
        raise ArgumentError("invalid pass for this image size")
    return ny.value, nx.value
