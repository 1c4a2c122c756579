"""The oracle's dewarp! (src/dewarp.jl:233-248) pinned without restating it:
the tensor-product natural cubic spline is evaluated independently by nested
scipy natural CubicSplines (first along dim 1 at the sample row, then along
dim 2), the identity map returns the frame itself, an integer-shift map moves
pixels exactly, out-of-domain and masked nodes are zero, and the Float32 path
rounds coordinates to Float32 as itp(T(rows), T(cols)) does."""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline


def nested_natural_spline(img, r, c):
    nr, nc = img.shape
    if not (1 <= r <= nr and 1 <= c <= nc):
        return 0.0
    col_vals = np.array([CubicSpline(np.arange(1, nr + 1), img[:, j], bc_type="natural")(r) for j in range(nc)])
    return float(CubicSpline(np.arange(1, nc + 1), col_vals, bc_type="natural")(c))


def test_dewarp_equals_nested_natural_splines(orc):
    rng = np.random.default_rng(3)
    img = rng.random((17, 23))
    rows = rng.uniform(0.5, 17.5, (6, 7))
    cols = rng.uniform(0.5, 23.5, (6, 7))
    rows[0, 0], cols[0, 0] = 1.0, 1.0     # the domain corners are inside
    rows[1, 0], cols[1, 0] = 17.0, 23.0
    got = orc.dewarp(img, rows, cols)
    ref = np.array([[nested_natural_spline(img, rows[i, j], cols[i, j]) for j in range(7)] for i in range(6)])
    assert np.allclose(got, ref, rtol=0, atol=1e-12)
    assert np.any(ref == 0.0) and np.any(ref != 0.0)  # both in- and out-of-domain samples occurred


def test_dewarp_identity_shift_and_mask(orc):
    rng = np.random.default_rng(4)
    img = rng.random((30, 40))
    I, J = np.meshgrid(np.arange(1.0, 31.0), np.arange(1.0, 41.0), indexing="ij")
    assert np.allclose(orc.dewarp(img, I, J), img, rtol=0, atol=1e-12)
    sh = orc.dewarp(img, I[:20, :30] + 3, J[:20, :30] + 5)
    assert np.allclose(sh, img[3:23, 5:35], rtol=0, atol=1e-12)
    m = np.zeros((20, 30), bool)
    m[4:9, 2:20] = True
    shm = orc.dewarp(img, I[:20, :30] + 3, J[:20, :30] + 5, mask=m)
    assert np.all(shm[m] == 0) and np.array_equal(shm[~m], sh[~m])


def test_dewarp_float32_rounds_coordinates(orc):
    rng = np.random.default_rng(5)
    img = rng.random((64, 64))
    rows = np.full((1, 3), 33.0) + np.array([[1e-9, 3e-6, 0.25]])
    cols = np.full((1, 3), 40.5)
    g32 = orc.dewarp(img.astype(np.float32), rows, cols, prec="f32")
    g64 = orc.dewarp(img, rows, cols)
    r32 = orc.dewarp(img, rows.astype(np.float32).astype(np.float64), cols)  # T(rows) in FP64 arithmetic
    assert np.allclose(g32, r32, rtol=0, atol=2e-6)
    assert np.allclose(g32, g64, rtol=0, atol=1e-4)
