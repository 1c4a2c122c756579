"""The reference's real-data test on the oracle: PIV Challenge case A
(test/test_reference.jl:8-26), a tip vortex with seeding loss in its core,
1024 x 1280 8-bit pair (fixture written by tests/golden/make_piv_challenge_a.py).
A 64 -> 32 padded, Gaussian-apodized multipass run gives finite u, v
everywhere, valid displacements under 16 px and fewer than a quarter of the
nodes flagged.
"""
import pathlib

import numpy as np

GOLDEN = pathlib.Path(__file__).parent / "golden" / "piv_challenge_a.npz"


def case_a():
    z = np.load(GOLDEN)
    A = np.asfortranarray(z["A"].astype(np.float64) / 255.0)
    B = np.asfortranarray(z["B"].astype(np.float64) / 255.0)
    return A, B


def check_case_a_bands(r):
    assert np.all(np.isfinite(r["u"])) and np.all(np.isfinite(r["v"]))
    ok = ~r["outliers"]
    assert np.hypot(r["u"][ok], r["v"][ok]).max() < 16
    assert r["outliers"].sum() / r["outliers"].size < 0.25


def test_piv_challenge_a_oracle(orc, hh):
    A, B = case_a()
    assert A.shape == B.shape == (1024, 1280)
    passes = hh.multipass_parameters([64, 32], padding=True, apodization="gauss")
    r = orc.run_piv(A, B, [p.to_struct() for p in passes])
    check_case_a_bands(r)
