"""Error behaviour of the C ABI on inputs that must never reach a kernel:
non-finite predictor or node values (they would become NaN sample
positions), non-finite knots, and job grid indices below 1. Each is
rejected before any device work with the reference's error classes
(ArgumentError for values, DimensionMismatch for shapes), and the context
stays usable afterwards."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair(hh):
    return hh.synthetic.particle_pair(128, 128, hh.synthetic.uniform_flow(1.0, 0.5), seed=4)


def test_nonfinite_predictor_rejected(hh, ctx, pair):
    ctx.set_pair(pair.A, pair.B)
    ctx.prefilter()
    py = np.linspace(16.5, 112.5, 7)
    px = np.linspace(16.5, 112.5, 7)
    pu = np.zeros((7, 7), np.float32)
    pv = np.zeros((7, 7), np.float32)
    for bad in (np.nan, np.inf, -np.inf):
        q = pu.copy()
        q[3, 2] = bad
        with pytest.raises(hh.ArgumentError):
            ctx.deform(py, px, q, pv)
        with pytest.raises(hh.ArgumentError):
            ctx.deform(py, px, pu, q)
    ky = py.copy()
    ky[0] = np.nan
    with pytest.raises(hh.ArgumentError):
        ctx.deform(ky, px, pu, pv)
    ctx.deform(py, px, pu, pv)  # still usable; zero predictor = identity
    wA, wB = ctx.images()
    assert np.allclose(wA, pair.A, atol=2e-5 * np.abs(pair.A).max())


def test_nonfinite_nodes_rejected_by_energy(hh, ctx, pair):
    ctx.set_pair(pair.A, pair.B)
    p = hh.PIVParameters(window_size=16, overlap=8, apodization="gauss")
    py = np.linspace(16.5, 112.5, 7)
    px = np.linspace(16.5, 112.5, 7)
    pu = np.zeros((7, 7), np.float32)
    bad = pu.copy()
    bad[0, 0] = np.nan
    with pytest.raises(hh.ArgumentError):
        ctx.grad_energy(p, py, px, bad, pu)
    with pytest.raises(hh.ArgumentError):
        ctx.energy(p, py, px, pu, bad)
    with pytest.raises(hh.ArgumentError):
        ctx.refine_nodes(p, py, px, bad, pu, niter=1)
    E, gu, gv = ctx.grad_energy(p, py, px, pu, pu)
    assert np.isfinite(E) and np.all(np.isfinite(gu)) and np.all(np.isfinite(gv))


def test_planes_reject_bad_jobs(hh, ctx, pair):
    ctx.set_pair(pair.A, pair.B)
    p = hh.PIVParameters(window_size=16, overlap=8, padding=True, apodization="gauss")
    x, y, jobs = hh.pass_grid(p, 128, 128)
    for row in (0, 1):
        bad = jobs[:, :4].copy()
        bad[row, 2] = 0
        with pytest.raises(hh.DimensionMismatch):
            ctx.correlation_planes(p, bad)
    pl = ctx.correlation_planes(p, jobs[:, :4])
    assert pl.shape == (32, 32, 4) and np.all(np.isfinite(pl))
