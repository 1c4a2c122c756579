"""Stereo per-camera analysis on dewarped images (src/dewarp.jl,
src/stereo.jl) for the mi355x backend.

``ImageDewarper`` holds what the reference's dewarper precomputes per camera
(src/dewarp.jl:196-221): the source pixel coordinate of every world-grid node
and the out-of-view mask. Building it needs the camera calibration
(``world_to_pixel``), which is the reference's host-side code; here it is
built from the coordinate map itself. ``dewarp`` resamples a frame on the
GPU (dewarp!, src/dewarp.jl:233-248). ``run_piv_stereo`` is the per-camera
part of the reference's stereo driver (_run_piv_stereo!,
src/stereo.jl:186-192): each camera's pair is dewarped on the device and
analyzed by run_piv there with the union mask, without a host round trip.
The (u, v, w) least-squares reconstruction (reconstruct_stereo) works on the
vector grid with the calibrations in Float64 on the host; a caller passes it
as ``reconstruct``.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

from . import _lib
from .params import ArgumentError, DimensionMismatch, PIVParameters, pass_grid_counts
from .pipeline import PIVResult, PIVWorkspace, _check_mask, _resolve_backend


@dataclass
class ImageDewarper:
    """ImageDewarper (src/dewarp.jl:196-221) without its calibration: rows,
    cols (ny x nx float64, 1-based source coordinates), mask (True = out of
    view), the camera's image_size (rows, cols) and a grid tag standing in
    for the shared DewarpGrid."""

    rows: np.ndarray
    cols: np.ndarray
    mask: np.ndarray
    image_size: tuple
    grid: object = None

    @classmethod
    def from_map(cls, rows, cols, image_size, grid=None):
        """The constructor's view test (src/dewarp.jl:205-215): a node whose
        source coordinate is non-finite or outside the frame is out of view
        (coordinates 0, mask True)."""
        nr, nc = int(image_size[0]), int(image_size[1])
        if nr < 1 or nc < 1:
            raise ArgumentError(f"image_size must be positive, got {image_size}")
        r = np.array(rows, dtype=np.float64, order="F", copy=True)
        c = np.array(cols, dtype=np.float64, order="F", copy=True)
        if r.shape != c.shape:
            raise DimensionMismatch(f"rows {r.shape} and cols {c.shape} differ")
        with np.errstate(invalid="ignore"):
            inside = np.isfinite(r) & np.isfinite(c) & (r >= 1) & (r <= nr) & (c >= 1) & (c <= nc)
        r[~inside] = 0.0
        c[~inside] = 0.0
        return cls(r, c, ~inside, (nr, nc), grid)

    @property
    def grid_size(self):
        return self.rows.shape


def _load_map(ctx, dw: ImageDewarper):
    ctx.set_dewarp_map(dw.rows, dw.cols, dw.mask, dw.image_size)


def dewarp(dw: ImageDewarper, img, *, backend: str = "mi355x", workspace: Optional[PIVWorkspace] = None):
    """dewarp(dw, img) (src/dewarp.jl:250-256) on the GPU: a float32 matrix
    of the grid's size, 0 at out-of-view nodes."""
    be = _resolve_backend(backend)
    img = np.asarray(img)
    if img.shape != tuple(dw.image_size):
        raise DimensionMismatch(f"image size {img.shape} does not match the dewarper's {dw.image_size}")
    ws = workspace or PIVWorkspace(be)
    try:
        _load_map(ws.ctx, dw)
        return ws.ctx.dewarp(img)
    finally:
        if workspace is None:
            ws.close()


@dataclass
class StereoCameraResults:
    """The per-camera 2C results of a stereo pair on the dewarped images
    (the cam1 / cam2 fields of StereoPIVResult, src/stereo.jl:32-58) and the
    union mask both were analyzed with."""

    cam1: PIVResult
    cam2: PIVResult
    mask: np.ndarray


def run_piv_stereo(A1, B1, A2, B2, dw1: ImageDewarper, dw2: ImageDewarper, params, *, effort=None,
                   backend: str = "mi355x", mask=None, workspace: Optional[PIVWorkspace] = None,
                   reconstruct: Optional[Callable] = None, predictor_smoothing: bool = True, **kwargs):
    """run_piv_stereo(A1, B1, A2, B2, dw1, dw2, params; mask, backend)
    (src/stereo.jl:134-192): the backend's option scope is checked before any
    dewarp, the dewarpers must share one grid, the analysis mask is
    dw1.mask | dw2.mask (| mask), and each camera's pair is dewarped and
    analyzed with run_piv on the device. Returns ``reconstruct(cam1, cam2,
    mask)`` when given, else the StereoCameraResults."""
    if effort is not None:
        raise ArgumentError("effort cannot be combined with explicit PIVParameters or pass schedules")
    be = _resolve_backend(backend)
    passes = [params] if isinstance(params, PIVParameters) else list(params)
    if not passes:
        raise ArgumentError("at least one pass is required")
    unknown = set(kwargs) - {"threaded", "mask_threshold"}
    if unknown:
        raise ArgumentError(f"unsupported run_piv_stereo keyword(s): {', '.join(sorted(unknown))}")
    if dw1.grid_size != dw2.grid_size or dw1.grid != dw2.grid:
        raise ArgumentError("the two dewarpers must share the same DewarpGrid")
    node_mask = np.asarray(dw1.mask, bool) | np.asarray(dw2.mask, bool)
    if mask is not None:
        m = np.asarray(mask, bool)
        if m.shape != node_mask.shape:
            raise DimensionMismatch(f"mask size {m.shape} does not match the dewarp grid size {node_mask.shape}")
        node_mask = node_mask | m
    m, thr = _check_mask(node_mask, kwargs.get("mask_threshold", 0.5), node_mask.shape)
    ws = workspace or PIVWorkspace(be)
    try:
        ctx = ws.ctx
        ctx.check_passes(passes)  # _check_backend_params before the dewarps
        for p in passes:
            pass_grid_counts(p, *node_mask.shape)
        out = []
        for A, B, dw in ((A1, B1, dw1), (A2, B2, dw2)):
            A, B = np.asarray(A), np.asarray(B)
            if A.shape != tuple(dw.image_size) or B.shape != tuple(dw.image_size):
                raise DimensionMismatch(f"image size {A.shape} does not match the dewarper's {dw.image_size}")
            _load_map(ctx, dw)
            ctx.dewarp_pair(A, B)
            ctx.set_mask(m, thr)
            r = ctx.run_piv(passes, predictor_smoothing)
            out.append(PIVResult(r["x"], r["y"], r["u"], r["v"], r["peak_ratio"], r["correlation_moment"],
                                 r["uncertainty_u"], r["uncertainty_v"], r["outliers"], r["mask"], passes[-1]))
        ctx.set_mask(None)
    finally:
        if workspace is None:
            ws.close()
    res = StereoCameraResults(out[0], out[1], node_mask)
    return reconstruct(res.cam1, res.cam2, res.mask) if reconstruct is not None else res
