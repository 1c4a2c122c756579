"""PIVParameters and multipass_parameters: the host mirror of the reference's
pass configuration (src/types.jl:212-291, src/pipeline.jl:509-524), plus its
C-ABI form ``hh_pass`` (include/hammerhead_mi355x.h).

Validation errors raise :class:`ArgumentError`, the reference's exception for
invalid options; shape errors raise :class:`DimensionMismatch`.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field, replace
from typing import Sequence


class ArgumentError(ValueError):
    """Julia's ArgumentError: an invalid or unsupported option."""


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch: inconsistent array shapes."""


# ---- enumerations of include/hammerhead_mi355x.h --------------------------
METHODS = {"cross": 0, "phase": 1}
APODS = {"none": 0, "gauss": 1}
SUBPIX = {"gauss3": 0, "gauss9": 1, "gauss2d": 2, "none": 3}
FINDERS = {"regionalmax": 0, "exclusion": 1}
V_PEAK_RATIO, V_MOMENT, V_VELMAG, V_UOD = 1, 2, 3, 4
MAX_VALIDATORS = 4


class hh_validator(C.Structure):
    _fields_ = [("kind", C.c_int32), ("neighborhood", C.c_int32), ("a", C.c_double), ("b", C.c_double)]


class hh_pass(C.Structure):
    _fields_ = [
        ("window_size", C.c_int32 * 2),
        ("search_area_size", C.c_int32 * 2),
        ("overlap", C.c_int32 * 2),
        ("correlation_method", C.c_int32),
        ("padding", C.c_int32),
        ("apodization", C.c_int32),
        ("subpixel_method", C.c_int32),
        ("n_peaks", C.c_int32),
        ("peak_finder", C.c_int32),
        ("uncertainty", C.c_int32),
        ("uod_enable", C.c_int32),
        ("uod_threshold", C.c_double),
        ("uod_neighborhood", C.c_int32),
        ("replace_outliers", C.c_int32),
        ("min_peak_ratio", C.c_double),
        ("max_iterations", C.c_int32),
        ("keep_correlation_planes", C.c_int32),
        ("convergence_tol", C.c_double),
        ("n_validators", C.c_int32),
        ("reserved", C.c_int32),
        ("validators", hh_validator * MAX_VALIDATORS),
    ]


def _pair(x, name):
    if isinstance(x, bool):
        raise ArgumentError(f"{name} must be an Int or a 2-tuple of Ints")
    if isinstance(x, int):
        return (x, x)
    t = tuple(int(v) for v in x)
    if len(t) != 2:
        raise ArgumentError(f"{name} must be an Int or a 2-tuple of Ints, got {x}")
    return t


def parse_validator(spec):
    """parse_validator (src/quality.jl:437-458): ``("peak_ratio", 1.3)``,
    ``("correlation_moment", 0.5)``, ``("uod", {"threshold": 2.0, ...})``
    or ``("velocity_magnitude", {"max": 50})``."""
    if isinstance(spec, hh_validator):
        return spec
    try:
        name, val = spec
    except (TypeError, ValueError):
        raise ArgumentError(f"cannot interpret {spec!r} as a validator")
    if isinstance(val, (int, float)) and not isinstance(val, bool):
        if name == "peak_ratio":
            return hh_validator(V_PEAK_RATIO, 0, float(val), 0.0)
        if name == "correlation_moment":
            return hh_validator(V_MOMENT, 0, float(val), 0.0)
        raise ArgumentError(f"unknown validator :{name} (expected :peak_ratio or :correlation_moment)")
    if isinstance(val, dict):
        if name in ("uod", "universal_outlier"):
            return hh_validator(V_UOD, int(val.get("neighborhood_size", 2)), float(val["threshold"]),
                                float(val.get("epsilon", 0.1)))
        if name == "velocity_magnitude":
            return hh_validator(V_VELMAG, 0, float(val.get("min", 0.0)), float(val["max"]))
        raise ArgumentError(f"unknown validator :{name} (expected :uod or :velocity_magnitude)")
    raise ArgumentError(f"cannot interpret {spec!r} as a validator")


@dataclass(frozen=True)
class PIVParameters:
    """Validated configuration of one interrogation pass (src/types.jl:212-291).

    Keyword names, defaults and validation rules follow the reference; symbol
    values are plain strings (``"cross"``, ``"gauss"``, ...).
    """

    window_size: tuple = (32, 32)
    search_area_size: tuple | None = None
    overlap: tuple = (16, 16)
    correlation_method: str = "cross"
    padding: bool = False
    apodization: str = "none"
    subpixel_method: str = "gauss3"
    n_peaks: int = 3
    peak_finder: str = "regionalmax"
    uncertainty: bool = False
    uod_enable: bool = True
    uod_threshold: float = 2.0
    uod_neighborhood: int = 2
    min_peak_ratio: float = 1.0
    validation: tuple = ()
    replace_outliers: bool = True
    max_iterations: int = 1
    convergence_tol: float = 0.05
    keep_correlation_planes: bool = False

    def __post_init__(self):
        ws = _pair(self.window_size, "window_size")
        ss = ws if self.search_area_size is None else _pair(self.search_area_size, "search_area_size")
        ov = _pair(self.overlap, "overlap")
        object.__setattr__(self, "window_size", ws)
        object.__setattr__(self, "search_area_size", ss)
        object.__setattr__(self, "overlap", ov)
        if not all(w >= 4 for w in ws):
            raise ArgumentError(f"window_size must be at least 4 in each dimension, got {ws}")
        if not all(s >= w for s, w in zip(ss, ws)):
            raise ArgumentError(f"search_area_size must be at least window_size in each dimension, "
                                f"got search_area_size={ss} for window_size={ws}")
        if not all((s - w) % 2 == 0 for s, w in zip(ss, ws)):
            raise ArgumentError(f"search_area_size - window_size must be even in each dimension so the "
                                f"areas share a center, got search_area_size={ss} for window_size={ws}")
        if not all(0 <= o < w for o, w in zip(ov, ws)):
            raise ArgumentError(f"overlap must satisfy 0 ≤ overlap < window_size, got overlap={ov} "
                                f"for window_size={ws}")
        if self.correlation_method not in METHODS:
            raise ArgumentError(f"correlation_method must be :cross or :phase, got :{self.correlation_method}")
        if self.apodization not in APODS:
            raise ArgumentError(f"apodization must be :none or :gauss, got :{self.apodization}")
        if self.subpixel_method not in SUBPIX:
            raise ArgumentError("subpixel_method must be :gauss3, :gauss9, :gauss2d, or :none, "
                                f"got :{self.subpixel_method}")
        if not (isinstance(self.n_peaks, int) and self.n_peaks >= 1):
            raise ArgumentError(f"n_peaks must be at least 1, got {self.n_peaks}")
        if self.peak_finder not in FINDERS:
            raise ArgumentError(f"peak_finder must be :exclusion or :regionalmax, got :{self.peak_finder}")
        if not self.uod_threshold > 0:
            raise ArgumentError(f"uod_threshold must be positive, got {self.uod_threshold}")
        if not self.uod_neighborhood >= 1:
            raise ArgumentError(f"uod_neighborhood must be at least 1, got {self.uod_neighborhood}")
        if not self.min_peak_ratio >= 0:
            raise ArgumentError(f"min_peak_ratio must be non-negative, got {self.min_peak_ratio}")
        if not self.max_iterations >= 1:
            raise ArgumentError(f"max_iterations must be at least 1, got {self.max_iterations}")
        if not self.convergence_tol >= 0:
            raise ArgumentError(f"convergence_tol must be non-negative, got {self.convergence_tol}")
        vals = tuple(parse_validator(v) for v in self.validation)
        if len(vals) > MAX_VALIDATORS:
            raise ArgumentError(f"at most {MAX_VALIDATORS} extra validators are supported")
        object.__setattr__(self, "validation", vals)
        object.__setattr__(self, "uod_threshold", float(self.uod_threshold))
        object.__setattr__(self, "min_peak_ratio", float(self.min_peak_ratio))
        object.__setattr__(self, "convergence_tol", float(self.convergence_tol))

    def to_struct(self) -> hh_pass:
        """The C-ABI ``hh_pass`` of this pass."""
        p = hh_pass()
        p.window_size[:] = self.window_size
        p.search_area_size[:] = self.search_area_size
        p.overlap[:] = self.overlap
        p.correlation_method = METHODS[self.correlation_method]
        p.padding = int(self.padding)
        p.apodization = APODS[self.apodization]
        p.subpixel_method = SUBPIX[self.subpixel_method]
        p.n_peaks = self.n_peaks
        p.peak_finder = FINDERS[self.peak_finder]
        p.uncertainty = int(self.uncertainty)
        p.uod_enable = int(self.uod_enable)
        p.uod_threshold = self.uod_threshold
        p.uod_neighborhood = self.uod_neighborhood
        p.replace_outliers = int(self.replace_outliers)
        p.min_peak_ratio = self.min_peak_ratio
        p.max_iterations = self.max_iterations
        p.keep_correlation_planes = int(self.keep_correlation_planes)
        p.convergence_tol = self.convergence_tol
        p.n_validators = len(self.validation)
        for i, v in enumerate(self.validation):
            p.validators[i] = v
        return p

    def replace(self, **kw) -> "PIVParameters":
        return replace(self, **kw)


def multipass_parameters(window_sizes: Sequence, overlap_fraction: float = 0.5, final: dict | None = None,
                         **kwargs) -> list:
    """multipass_parameters (src/pipeline.jl:509-524): one pass per window
    size with overlap = floor(window * overlap_fraction); ``final`` overrides
    apply to the last pass only."""
    if not (0 <= overlap_fraction < 1):
        raise ArgumentError(f"overlap_fraction must be in [0, 1), got {overlap_fraction}")
    if len(window_sizes) == 0:
        raise ArgumentError("window_sizes must not be empty")
    final = final or {}
    n = len(window_sizes)
    out = []
    for i, w in enumerate(window_sizes):
        ws = _pair(w, "window_size")
        extra = dict(kwargs)
        if i == n - 1:
            extra.update(final)
        ov = tuple(int(math.floor(s * overlap_fraction)) for s in ws)
        # Julia keyword splatting: later keywords (kwargs, final) override
        kw = dict(window_size=ws, overlap=ov)
        kw.update(extra)
        out.append(PIVParameters(**kw))
    return out


def pass_grid_counts(p: PIVParameters, nr: int, nc: int):
    """Row/column start ranges of pass_grid (src/pipeline.jl:529-555)."""
    wr, wc = p.window_size
    sr, sc = p.search_area_size
    mr, mc = (sr - wr) // 2, (sc - wc) // 2
    if sr > nr or sc > nc:
        raise ArgumentError(f"search area {p.search_area_size} exceeds image size {(nr, nc)}")
    rows = range(1 + mr, nr - wr - mr + 2, wr - p.overlap[0])
    cols = range(1 + mc, nc - wc - mc + 2, wc - p.overlap[1])
    return rows, cols
