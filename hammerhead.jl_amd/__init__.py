"""hammerhead.jl_amd — MI355X-native (gfx950 HIP) backend for Hammerhead.jl's
WIDIM multi-pass correlation path, with a Python host mirror of the
reference's run_piv / multipass_parameters / PIVParameters / PIVResult
interface (src/pipeline.jl:361-524, src/types.jl:212-373).

Import name: the directory is loaded as ``hammerhead_amd`` (see
``__graft_entry__.load_package``).
"""
from .params import (ArgumentError, DimensionMismatch, PIVParameters, hh_pass, hh_validator,
                     multipass_parameters, pass_grid_counts)
from . import synthetic
from ._lib import Context, DeviceError, HH_FORM_FFT32, HH_FORM_MFMA16, HH_SRC_RAW, HH_SRC_WARP, grid_size
from .pipeline import (PIVResult, PIVWorkspace, pass_grid, piv_workspace, run_piv, run_piv_ensemble,
                       run_piv_sequence)
from .stereo import ImageDewarper, StereoCameraResults, dewarp, run_piv_stereo

__all__ = [
    "ArgumentError", "DimensionMismatch", "PIVParameters", "multipass_parameters", "run_piv", "run_piv_ensemble",
    "run_piv_sequence",
    "PIVResult",
    "PIVWorkspace", "piv_workspace", "pass_grid", "Context", "DeviceError", "synthetic",
    "ImageDewarper", "StereoCameraResults", "dewarp", "run_piv_stereo",
]
