import os
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def hh():
    import __graft_entry__ as ge

    return ge.load_package()


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ctx(hh):
    c = hh.Context(0)
    yield c
    c.close()


def add_particle(a, centroid, diameter):
    """The reference tests' particle helper (test/runtests.jl:11-19):
    sigma = diameter/2, cut at 5 sigma, peak 1, 1-based (row, col)."""
    import numpy as np

    sigma = diameter / 2
    nr, nc = a.shape
    i = np.arange(1, nr + 1)[:, None]
    j = np.arange(1, nc + 1)[None, :]
    d2 = (i - centroid[0]) ** 2 + (j - centroid[1]) ** 2
    a += np.where(d2 > (5 * sigma) ** 2, 0.0, np.exp(-0.5 * d2 / sigma ** 2))
    return a


def particle_pair(shape, positions, dv, du, diameter=3.0):
    """particle_pair (test/runtests.jl:26-34)."""
    import numpy as np

    A = np.zeros(shape)
    B = np.zeros(shape)
    for p in positions:
        add_particle(A, p, diameter)
        add_particle(B, (p[0] + dv, p[1] + du), diameter)
    return np.asfortranarray(A), np.asfortranarray(B)


def low_snr_pairs(n=96, npairs=24, du=1.5, dv=-2.0, nparticles=120, noise=1.5, seed=31):
    """The ensemble scene of test/test_ensemble.jl:7-17 regenerated with
    numpy's RNG: 120 particles over [-10, n+10]^2, a uniform shift, and heavy
    uniform noise 1.5 U[0, 1] added to each frame."""
    import numpy as np

    rng = np.random.default_rng(seed)
    pairs = []
    for _ in range(npairs):
        pos = [(rng.random() * (n + 20) - 10, rng.random() * (n + 20) - 10) for _ in range(nparticles)]
        A, B = particle_pair((n, n), pos, dv, du)
        A = A + noise * rng.random((n, n))
        B = B + noise * rng.random((n, n))
        pairs.append((np.asfortranarray(A), np.asfortranarray(B)))
    return pairs
