"""Device dewarping and the stereo per-camera path against the oracle.

* hh_dewarp (dewarp!, src/dewarp.jl:233-248) vs the FP64 oracle on a smooth
  projective map with out-of-view nodes: within 2e-5 of the frame's range at
  every node (the device resolves source positions to 2^-16 px and computes
  in float32), out-of-view nodes exactly 0; the identity map returns the
  frame; an integer shift moves pixels.
* run_piv_stereo's per-camera analyses (src/stereo.jl:186-192): each camera
  dewarped on the device and analyzed by run_piv with the union mask gives
  the same field as run_piv on the host-fetched dewarped pair, and matches
  the FP64 oracle's run_piv on the oracle's dewarped pair to the parity bars
  (tests/parity_bars.py).
"""
import numpy as np
import pytest

from parity_bars import field_parity, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rig(hh):
    (r1, c1), (r2, c2), size = hh.synthetic.stereo_maps()
    dw1 = hh.ImageDewarper.from_map(r1, c1, size, grid="world")
    dw2 = hh.ImageDewarper.from_map(r2, c2, size, grid="world")
    (A1, B1), (A2, B2) = hh.synthetic.stereo_frames(size)
    return dw1, dw2, (A1, B1), (A2, B2)


def test_dewarp_vs_fp64_oracle(hh, orc, ctx, rig):
    dw1, dw2, (A1, B1), (A2, B2) = rig
    for dw, img in ((dw1, A1), (dw2, B2)):
        ctx.set_dewarp_map(dw.rows, dw.cols, dw.mask, dw.image_size)
        g = ctx.dewarp(img)
        o = orc.dewarp(img.astype(np.float64), dw.rows, dw.cols, dw.mask)
        err = np.abs(g - o).max() / np.abs(img).max()
        say("dewarp max err / range", err, "out of view", int(dw.mask.sum()))
        assert err < 2e-5
        assert np.all(g[dw.mask] == 0)
        assert dw.mask.any()


def test_dewarp_identity_and_shift(hh, ctx):
    rng = np.random.default_rng(9)
    img = rng.random((96, 80)).astype(np.float32)
    I, J = np.meshgrid(np.arange(1.0, 97.0), np.arange(1.0, 81.0), indexing="ij")
    ctx.set_dewarp_map(I, J, None, img.shape)
    assert np.allclose(ctx.dewarp(img), img, rtol=0, atol=2e-6)
    ctx.set_dewarp_map(I[:60, :50] + 7, J[:60, :50] + 11, None, img.shape)
    assert np.allclose(ctx.dewarp(img), img[7:67, 11:61], rtol=0, atol=2e-6)
    with pytest.raises(hh.DimensionMismatch):
        ctx.dewarp(img[:90])


def test_stereo_cameras_vs_oracle(hh, orc, rig):
    dw1, dw2, (A1, B1), (A2, B2) = rig
    passes = hh.multipass_parameters([32, 16], padding=True, apodization="gauss",
                                     final=dict(overlap=8, max_iterations=2, convergence_tol=0.0, uncertainty=True))
    say("stereo: device dewarp + run_piv per camera")
    res = hh.run_piv_stereo(A1, B1, A2, B2, dw1, dw2, passes)
    node_mask = dw1.mask | dw2.mask
    assert np.array_equal(res.mask, node_mask)
    for k, (dw, (A, B), cam) in enumerate(((dw1, (A1, B1), res.cam1), (dw2, (A2, B2), res.cam2))):
        with hh.Context(0) as c:
            c.set_dewarp_map(dw.rows, dw.cols, dw.mask, dw.image_size)
            a, b = c.dewarp(A), c.dewarp(B)
        # forwarding: the same as run_piv on the dewarped pair fetched to the host
        ref = hh.run_piv(a, b, passes, mask=node_mask)
        for f in ("u", "v", "peak_ratio", "correlation_moment", "uncertainty_u", "outliers", "mask"):
            assert np.array_equal(getattr(cam, f), getattr(ref, f), equal_nan=True), (k, f)
        assert cam.mask.any()
        # parity: the FP64 oracle dewarps and analyzes the same frames
        oa = orc.dewarp(A.astype(np.float64), dw.rows, dw.cols, dw.mask)
        ob = orc.dewarp(B.astype(np.float64), dw.rows, dw.cols, dw.mask)
        o = orc.run_piv(oa, ob, [p.to_struct() for p in passes], mask=node_mask, uq_pert=True)
        g = {f: getattr(cam, f) for f in ("u", "v", "peak_ratio", "correlation_moment", "uncertainty_u",
                                          "uncertainty_v", "outliers", "mask")}
        field_parity(g, o, label=f"stereo camera {k + 1}", rms_tol=5e-3)
