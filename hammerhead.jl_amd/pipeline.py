"""run_piv and the backend seam of the host mirror.

``run_piv(imgA, imgB, passes; backend="mi355x")`` mirrors the reference's
driver (src/pipeline.jl:361-480) and returns a :class:`PIVResult` with the
reference's fields (src/types.jl:359-373). The whole schedule — prefilter,
deformation, correlation, validation, predictor smoothing and uncertainty —
runs on the GPU through ``hh_run_piv``. There is no CPU fallback: backend
``"cpu"`` is the reference itself and is not provided here.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .params import ArgumentError, DimensionMismatch, PIVParameters, pass_grid_counts


@dataclass
class PIVResult:
    """PIVResult{Float32} (src/types.jl:359-373)."""

    x: np.ndarray
    y: np.ndarray
    u: np.ndarray
    v: np.ndarray
    peak_ratio: np.ndarray
    correlation_moment: np.ndarray
    uncertainty_u: np.ndarray
    uncertainty_v: np.ndarray
    outliers: np.ndarray
    mask: np.ndarray
    parameters: PIVParameters
    correlation_planes: Optional[np.ndarray] = None
    scale: object = None

    def __repr__(self):
        ny, nx = self.u.shape
        return f"PIVResult{{Float32}}({nx}×{ny} grid, {int(self.outliers.sum())} outliers)"


_BACKENDS = {"mi355x"}


def _resolve_backend(backend: str) -> str:
    """_resolve_backend(::Val{sym}) (src/types.jl:79-86)."""
    if backend in _BACKENDS:
        return backend
    raise ArgumentError(f"unsupported Hammerhead backend :{backend}; this package provides backend = :mi355x "
                        "(the gfx950 device path). The CPU path is the reference package itself.")


class PIVWorkspace:
    """piv_workspace(; backend) (src/pipeline.jl:21-47): one device context
    whose buffers, prefiltered coefficients and engine pool are reused across
    run_piv calls on equally sized pairs. Not for concurrent use."""

    def __init__(self, backend: str = "mi355x", device: int = 0):
        self.backend = _resolve_backend(backend)
        self.device = device
        self.ctx = _lib.Context(device)
        self._ctx2 = None  # second context of a pipelined run_piv_sequence

    def second_context(self):
        if self._ctx2 is None:
            self._ctx2 = _lib.Context(self.device)
        return self._ctx2

    def close(self):
        self.ctx.close()
        if self._ctx2 is not None:
            self._ctx2.close()
            self._ctx2 = None


def piv_workspace(backend: str = "mi355x", device: int = 0) -> PIVWorkspace:
    return PIVWorkspace(backend, device)


def pass_grid(p: PIVParameters, nr: int, nc: int):
    """pass_grid (src/pipeline.jl:529-555), unmasked: x, y, jobs (4 x n int32)."""
    rows, cols = pass_grid_counts(p, nr, nc)
    wr, wc = p.window_size
    x = np.array([cs + (wc - 1) / 2 for cs in cols], dtype=np.float32)
    y = np.array([rs + (wr - 1) / 2 for rs in rows], dtype=np.float32)
    jobs = np.array([(gi + 1, gj + 1, rs, cs) for gj, cs in enumerate(cols) for gi, rs in enumerate(rows)],
                    dtype=np.int32).reshape(-1, 4).T
    return x, y, np.asfortranarray(jobs)


def _check_mask(mask, mask_threshold, shape):
    """The mask keyword checks of run_piv (src/pipeline.jl:393-396): the
    mask's size first, then the threshold's range."""
    m = None
    if mask is not None:
        m = np.asarray(mask, dtype=bool)
        if m.shape != tuple(shape):
            raise DimensionMismatch(f"mask must have the same size as the images, got {m.shape}")
    thr = float(mask_threshold)
    if not (0 < thr <= 1):
        raise ArgumentError(f"mask_threshold must be in (0, 1], got {mask_threshold}")
    return m, thr


def run_piv(imgA, imgB, passes, *, effort=None, backend: str = "mi355x", predictor_smoothing: bool = True,
            mask=None, workspace: Optional[PIVWorkspace] = None, form: int = _lib.HH_FORM_FFT32, **kwargs):
    """run_piv(imgA, imgB, passes; backend) -> PIVResult (src/pipeline.jl:361-449).

    ``passes`` is a PIVParameters or a list of them. The computation runs in
    float32 on the GPU (the reference's Float32 pipeline). ``mask`` (an
    image-sized boolean matrix) and ``mask_threshold`` act as in the
    reference: windows whose masked fraction reaches the threshold give NaN
    vectors flagged in ``result.mask``; other windows use their valid pixels."""
    if effort is not None:
        raise ArgumentError("effort cannot be combined with explicit PIVParameters or pass schedules")
    unknown = set(kwargs) - {"threaded", "uncertainty_backend", "mask_threshold", "scale", "roi"}
    if unknown:
        raise ArgumentError(f"unsupported run_piv keyword(s): {', '.join(sorted(unknown))}")
    if kwargs.get("roi") is not None:
        raise ArgumentError("backend :mi355x does not support roi yet")
    be = _resolve_backend(backend)
    if workspace is not None and workspace.backend != be:
        raise ArgumentError(f"workspace backend {workspace.backend} does not match run backend {be}")
    if isinstance(passes, PIVParameters):
        passes = [passes]
    passes = list(passes)
    if not passes:
        raise ArgumentError("at least one pass is required")
    A = np.asarray(imgA)
    B = np.asarray(imgB)
    if A.shape != B.shape:
        raise DimensionMismatch(f"images must have the same size, got {A.shape} and {B.shape}")
    m, thr = _check_mask(mask, kwargs.get("mask_threshold", 0.5), A.shape)
    ws = workspace or PIVWorkspace(be)
    try:
        ctx = ws.ctx
        ctx.check_passes(passes)
        for p in passes:
            pass_grid_counts(p, *A.shape)
        if form != _lib.HH_FORM_FFT32:
            ctx.set_form(form)
        ctx.set_mask(m, thr)
        ctx.set_pair(A, B)
        r = ctx.run_piv(passes, predictor_smoothing)
    finally:
        if workspace is None:
            ws.close()
    return PIVResult(r["x"], r["y"], r["u"], r["v"], r["peak_ratio"], r["correlation_moment"], r["uncertainty_u"],
                     r["uncertainty_v"], r["outliers"], r["mask"], passes[-1], None, kwargs.get("scale"))


def run_piv_ensemble(pairs: Sequence, params, *, effort=None, backend: str = "mi355x", predictor_smoothing: bool = True,
                     mask=None, workspace: Optional[PIVWorkspace] = None, **kwargs) -> PIVResult:
    """run_piv_ensemble(pairs, params; backend) -> PIVResult (src/ensemble.jl:47-84).

    ``pairs`` is a sequence of (imgA, imgB) arrays of one size; ``params`` a
    PIVParameters or a schedule. Every pass sums each window's correlation
    planes over all pairs in FP64 on the GPU and locates the peak once on the
    sum; with a schedule, the previous pass's ensemble field is the shared
    deformation predictor of every pair (replacement is forced on all but the
    last pass). With ``uncertainty`` on the last pass, the Wieneke statistics
    are pooled over the pairs. ``mask`` and ``mask_threshold`` act as in
    run_piv (src/ensemble.jl:151-154,254-264)."""
    if effort is not None:
        raise ArgumentError("effort cannot be combined with explicit PIVParameters or pass schedules")
    unknown = set(kwargs) - {"threaded", "mask_threshold", "scale", "progress", "preprocess", "image_type"}
    if unknown:
        raise ArgumentError(f"unsupported run_piv_ensemble keyword(s): {', '.join(sorted(unknown))}")
    be = _resolve_backend(backend)
    passes = [params] if isinstance(params, PIVParameters) else list(params)
    if not passes:
        raise ArgumentError("at least one pass is required")
    pairs = list(pairs)
    if not pairs:
        raise ArgumentError("pairs must not be empty")
    preprocess = kwargs.get("preprocess")
    arrays = []
    for A, B in pairs:
        A, B = np.asarray(A), np.asarray(B)
        if preprocess is not None:
            A, B = preprocess(A), preprocess(B)
        if A.shape != B.shape:
            raise DimensionMismatch(f"images must have the same size, got {A.shape} and {B.shape}")
        if arrays and A.shape != arrays[0][0].shape:
            raise DimensionMismatch(f"all pairs must share the image size {arrays[0][0].shape}, got {A.shape}")
        arrays.append((A, B))
    shape = arrays[0][0].shape
    m, thr = _check_mask(mask, kwargs.get("mask_threshold", 0.5), shape)
    ws = workspace or PIVWorkspace(be)
    try:
        ctx = ws.ctx
        ctx.check_passes(passes)
        for p in passes:
            pass_grid_counts(p, *shape)
        ctx.set_mask(m, thr)
        r = ctx.run_piv_ensemble(arrays, passes, predictor_smoothing)
    finally:
        if workspace is None:
            ws.close()
    return PIVResult(r["x"], r["y"], r["u"], r["v"], r["peak_ratio"], r["correlation_moment"], r["uncertainty_u"],
                     r["uncertainty_v"], r["outliers"], r["mask"], passes[-1], None, kwargs.get("scale"))


def run_piv_sequence(pairs: Sequence, params, *, effort=None, backend: str = "mi355x", preprocess=None,
                     output=None, progress=False, on_result=None, image_type=np.float32, mask=None, scale=None,
                     predictor_smoothing: bool = True, batch: Optional[int] = None,
                     workspace: Optional[PIVWorkspace] = None, form: int = _lib.HH_FORM_FFT32,
                     **kwargs) -> List[PIVResult]:
    """run_piv_sequence(pairs, params; backend) -> Vector{PIVResult}
    (src/io.jl:229-247, driver src/io.jl:307-358).

    ``pairs`` holds (imgA, imgB) matrices of any sizes (pairs may differ).
    Instead of one run_piv per pair, up to ``batch`` pairs (default: all) are
    packed into one device-resident atlas and run together: every sweep of
    every pass is one launch per stage over all pairs of the batch, and with
    fixed sweep counts the schedule replays as a hipGraph. Each pair's result
    equals run_piv on that pair alone. With more than one batch, two device
    contexts alternate: batch i + 1 is uploaded (and its schedule enqueued)
    while batch i still runs, and batch i's results are fetched after that.
    ``on_result(i, result)`` (1-based i) and a callable ``progress(i, n)`` run
    in pair order after each batch.
    Frames given as file paths and ``output`` (JLD2 persistence) belong to the
    reference's I/O layer and are rejected."""
    if effort is not None:
        raise ArgumentError("effort cannot be combined with explicit PIVParameters or pass schedules")
    unknown = set(kwargs) - {"threaded", "uncertainty_backend", "mask_threshold"}
    if unknown:
        raise ArgumentError(f"unsupported run_piv_sequence keyword(s): {', '.join(sorted(unknown))}")
    if output is not None:
        raise ArgumentError("backend :mi355x returns results in memory; output files are not supported")
    be = _resolve_backend(backend)
    passes = [params] if isinstance(params, PIVParameters) else list(params)
    if not passes:
        raise ArgumentError("at least one pass is required")
    pairs = list(pairs)
    if not pairs:
        raise ArgumentError("pairs must not be empty")
    m = None
    thr = kwargs.get("mask_threshold", 0.5)
    frames = []
    for A, B in pairs:
        if isinstance(A, (str, bytes)) or isinstance(B, (str, bytes)):
            raise ArgumentError("backend :mi355x takes in-memory image matrices; load frames first")
        A, B = np.asarray(A), np.asarray(B)
        if preprocess is not None:
            A, B = preprocess(A), preprocess(B)
        if A.shape != B.shape:
            raise DimensionMismatch(f"images must have the same size, got {A.shape} and {B.shape}")
        for p in passes:
            pass_grid_counts(p, *A.shape)
        m, thr = _check_mask(mask, thr, A.shape)
        frames.append((A, B))
    n = len(frames)
    step = n if batch is None else max(1, int(batch))
    ws = workspace or PIVWorkspace(be)
    results: List[PIVResult] = []
    try:
        ctxs = [ws.ctx] if step >= n else [ws.ctx, ws.second_context()]
        for ctx in ctxs:
            ctx.check_passes(passes)
            ctx.set_form(form)
            ctx.set_mask(m, thr)

        def emit(ctx, q0):
            for q in range(len(ctx.batch_shapes)):
                r = ctx.fetch_pair(q)
                res = PIVResult(r["x"], r["y"], r["u"], r["v"], r["peak_ratio"], r["correlation_moment"],
                                r["uncertainty_u"], r["uncertainty_v"], r["outliers"], r["mask"], passes[-1], None,
                                scale)
                results.append(res)
                i = q0 + q + 1
                if on_result is not None:
                    on_result(i, res)
                if callable(progress):
                    progress(i, n)

        pending = None
        for b, q0 in enumerate(range(0, n, step)):
            ctx = ctxs[b % len(ctxs)]
            ctx.set_pairs(frames[q0:q0 + step])
            ctx.run_piv_batch(passes, predictor_smoothing, fetch=False)
            if pending is not None:
                emit(*pending)
            pending = (ctx, q0)
        emit(*pending)
    finally:
        if workspace is None:
            ws.close()
    return results
