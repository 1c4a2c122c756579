"""PIV Challenge case A (test/test_reference.jl:8-26) on the device through
the C ABI: the reference's bands (finite field, valid |d| < 16 px, under a
quarter flagged) and parity with the oracle on the same real pair (flags
identical, u/v within 1e-3 px at every node, peak ratio and moment within 1 %),
in both forms: FP32 transforms, and the split-FP16
matrix-core kernels (the 32-px pass on k_corr_mfma64).
"""
import sys

import numpy as np
import pytest

from parity_bars import field_parity
from test_oracle_reference_images import case_a, check_case_a_bands

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", ["fft32", "mfma16"])
def test_piv_challenge_a_device_vs_oracle(hh, orc, form):
    A, B = case_a()
    passes = hh.multipass_parameters([64, 32], padding=True, apodization="gauss")
    print("piv challenge A", A.shape, form, flush=True)
    sys.stdout.flush()
    with hh.Context(0) as c:
        c.set_form(hh.HH_FORM_MFMA16 if form == "mfma16" else hh.HH_FORM_FFT32)
        c.set_pair(A.astype(np.float32), B.astype(np.float32))
        g = c.run_piv(passes)
    o = orc.run_piv(A, B, [p.to_struct() for p in passes])
    check_case_a_bands(g)
    assert np.array_equal(g["x"], o["x"]) and np.array_equal(g["y"], o["y"])
    field_parity(g, o, label=f"case A {form}", rms_tol=5e-3, uncertainty=False)
