"""Seeded synthetic particle-image pairs with known displacement.

Particle convention of the reference (src/synthetic_data.jl:170-189,
206-232): a Gaussian particle of 1/e^2 diameter d has sigma = d/4 and is
accumulated within round(c -/+ ceil(3 sigma)) in each axis (Julia `round`,
half to even), clipped to the image; frame B holds each particle displaced
by one forward-Euler step of the flow at its frame-A position
(src/synthetic_data.jl:145-159); additive Gaussian noise is then clamped at
zero (src/synthetic_data.jl:227-230). Coordinates are 1-based pixel
centres, x along columns and y along rows. Images are returned column-major
(Fortran order), like the reference's Julia matrices.

Positions are uniform over [-10, N+10] in each axis so the borders stay
seeded, with round(ppp * (N+20)^2) particles and peak intensity U[0.5, 1].
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable

import numpy as np


# ---- flows: (x, y) -> (u, v) displacement in pixels ------------------------

def uniform_flow(u0: float, v0: float) -> Callable:
    return lambda x, y: (np.full_like(x, u0), np.full_like(y, v0))


def lamb_oseen_flow(xc: float, yc: float, core_radius: float, peak_displacement: float) -> Callable:
    """Lamb-Oseen vortex: u_theta = (G/2 pi r)(1 - exp(-r^2/rc^2)); its
    maximum, at r = 1.12091 rc, equals peak_displacement."""
    xm = 1.1209064227313  # argmax of (1 - exp(-s^2))/s
    amp = peak_displacement / ((1 - math.exp(-xm * xm)) / xm)  # = G / (2 pi rc)

    def f(x, y):
        dx, dy = x - xc, y - yc
        r = np.hypot(dx, dy)
        s = r / core_radius
        with np.errstate(invalid="ignore", divide="ignore"):
            ut = np.where(s > 0, amp * (1 - np.exp(-s * s)) / np.where(s > 0, s, 1), 0.0)
            cu = np.where(r > 0, -dy / np.where(r > 0, r, 1), 0.0)
            cv = np.where(r > 0, dx / np.where(r > 0, r, 1), 0.0)
        return ut * cu, ut * cv

    return f


def solid_body_flow(xc: float, yc: float, omega: float, u0: float = 0.0, v0: float = 0.0) -> Callable:
    return lambda x, y: (-omega * (y - yc) + u0, omega * (x - xc) + v0)


def sine_shear_flow(amplitude: float, wavelength: float) -> Callable:
    return lambda x, y: (amplitude * np.sin(2 * np.pi * y / wavelength), np.zeros_like(y))


def linear_shear_flow(a: float, y0: float) -> Callable:
    return lambda x, y: (a * (y - y0), np.zeros_like(y))


# ---- rendering --------------------------------------------------------------

def render(nr: int, nc: int, xs, ys, intensity, diameter: float) -> np.ndarray:
    """Sum of Gaussian particles (generate_gaussian_particle!,
    src/synthetic_data.jl:176-189) into an nr x nc float64 image."""
    sigma = diameter / 4.0
    rad = int(math.ceil(3 * sigma))
    img = np.zeros((nr, nc), dtype=np.float64)
    keep = (xs > -rad - 1) & (xs < nc + rad + 2) & (ys > -rad - 1) & (ys < nr + rad + 2)
    xs, ys, intensity = xs[keep], ys[keep], intensity[keep]
    j0 = np.rint(xs - rad).astype(np.int64)
    i0 = np.rint(ys - rad).astype(np.int64)
    span = 2 * rad + 1
    for dj in range(span):
        j = j0 + dj
        for di in range(span):
            i = i0 + di
            ok = (j >= 1) & (j <= nc) & (i >= 1) & (i <= nr)
            # the box is round(c - r):round(c + r); with integer r it always
            # spans exactly 2r + 1 pixels
            r2 = (j[ok] - xs[ok]) ** 2 + (i[ok] - ys[ok]) ** 2
            val = intensity[ok] * np.exp(-r2 / (2 * sigma * sigma))
            np.add.at(img, (i[ok] - 1, j[ok] - 1), val)
    return img


@dataclass
class SyntheticPair:
    A: np.ndarray  # nr x nc, Fortran order
    B: np.ndarray
    flow: Callable
    seed: int

    def truth(self, x, y):
        """Displacement at frame-A positions (x, y): the midpoint truth of a
        symmetric deformation is evaluated by the caller as needed."""
        return self.flow(np.asarray(x, float), np.asarray(y, float))


def particle_pair(nr: int, nc: int, flow: Callable, seed: int, ppp: float = 0.03, diameter: float = 2.5,
                  noise: float = 0.02, dtype=np.float32, margin: float = 10.0) -> SyntheticPair:
    rng = np.random.default_rng(seed)
    n = int(round(ppp * (nr + 2 * margin) * (nc + 2 * margin)))
    xs = rng.uniform(-margin, nc + margin, n)
    ys = rng.uniform(-margin, nr + margin, n)
    inten = rng.uniform(0.5, 1.0, n)
    du, dv = flow(xs, ys)
    A = render(nr, nc, xs, ys, inten, diameter)
    B = render(nr, nc, xs + du, ys + dv, inten, diameter)
    if noise > 0:
        A += noise * rng.standard_normal((nr, nc))
        B += noise * rng.standard_normal((nr, nc))
        np.maximum(A, 0.0, out=A)
        np.maximum(B, 0.0, out=B)
    return SyntheticPair(np.asfortranarray(A, dtype=dtype), np.asfortranarray(B, dtype=dtype), flow, seed)


# ---- the workloads this backend is measured and checked on -----------------

def config_uniform_256(seed: int = 101) -> SyntheticPair:
    """256 x 256 Float64 pair, 0.02 ppp, d = 2.5, noise 0.01, uniform
    (u, v) = (2.3, -1.7)."""
    return particle_pair(256, 256, uniform_flow(2.3, -1.7), seed, ppp=0.02, noise=0.01, dtype=np.float64)


def config_vortex_1024(seed: int = 102, n: int = 1024) -> SyntheticPair:
    """1024 x 1024 Float32 pair, 0.03 ppp, noise 0.02, Lamb-Oseen vortex of
    peak displacement 6 px and core radius 80 px at the image centre."""
    c = (n + 1) / 2.0
    return particle_pair(n, n, lamb_oseen_flow(c, c, 80.0 * n / 1024, 6.0), seed, ppp=0.03, noise=0.02)


def config_rotation_4mp(seed: int = 103, n: int = 2048) -> SyntheticPair:
    """2048 x 2048 Float32 pair, 0.03 ppp, noise 0.02: solid-body rotation
    reaching 8 px at the corner pixels plus a global U[-0.5, 0.5] shift."""
    rng = np.random.default_rng(seed + 7919)
    u0, v0 = rng.uniform(-0.5, 0.5, 2)
    c = (n + 1) / 2.0
    omega = 8.0 / ((c - 1) * math.sqrt(2))
    return particle_pair(n, n, solid_body_flow(c, c, omega, u0, v0), seed, ppp=0.03, noise=0.02)


def config_sine_shear_1024(seed: int = 105, n: int = 1024) -> SyntheticPair:
    """1024 x 1024 Float32 pair, 0.03 ppp: u = 3 sin(2 pi y / 256), v = 0."""
    return particle_pair(n, n, sine_shear_flow(3.0, 256.0), seed, ppp=0.03, noise=0.02)


BATCH_SIZES = ((512, 512), (768, 1024), (1024, 1280), (2048, 2048))  # (rows, cols)


def config_batch(npairs: int = 64, seed0: int = 104000, sizes=BATCH_SIZES, ppp: float = 0.03):
    """The batching workload: npairs Float32 pairs, pair i of size
    sizes[i % len(sizes)] and seed seed0 + i, alternating a Lamb-Oseen vortex
    (peak 8 px, core radius n/10) and a linear shear reaching 8 px at the
    image edges, plus a per-pair uniform shift U[-1, 1] px (all <= 10 px)."""
    out = []
    for i in range(npairs):
        nr, nc = sizes[i % len(sizes)]
        rng = np.random.default_rng(seed0 + i + 7919)
        u0, v0 = rng.uniform(-1.0, 1.0, 2)
        if i % 2 == 0:
            base = lamb_oseen_flow((nc + 1) / 2.0, (nr + 1) / 2.0, min(nr, nc) / 10.0, 8.0)
        else:
            base = linear_shear_flow(16.0 / nr, (nr + 1) / 2.0)

        def flow(x, y, base=base, u0=u0, v0=v0):
            u, v = base(x, y)
            return u + u0, v + v0

        out.append(particle_pair(nr, nc, flow, seed0 + i, ppp=ppp, noise=0.02))
    return out


def config_ensemble(npairs: int = 500, nr: int = 1024, nc: int = 1280, seed0: int = 104500, ppp: float = 0.005,
                    noise: float = 0.05):
    """The ensemble workload: npairs Float32 pairs of nr x nc at low seeding
    (0.005 ppp) and heavy noise (sigma 0.05), all of one steady Lamb-Oseen
    vortex (peak 6 px, core radius nc / 10) plus a uniform (1.2, -0.8) px
    shift; pair i is seeded seed0 + i. Returns (pairs, flow)."""
    base = lamb_oseen_flow((nc + 1) / 2.0, (nr + 1) / 2.0, nc / 10.0, 6.0)

    def flow(x, y):
        u, v = base(x, y)
        return u + 1.2, v - 0.8

    return [particle_pair(nr, nc, flow, seed0 + i, ppp=ppp, noise=noise) for i in range(npairs)], flow


def midpoint_truth(flow, x, y, iters: int = 8):
    """Displacement whose midpoint lies at (x, y): the truth a symmetric
    deformation measures (bench/run_benchmarks.jl:199-207)."""
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    x0, y0 = x.copy(), y.copy()
    for _ in range(iters):
        u, v = flow(x0, y0)
        x0 = x - 0.5 * u
        y0 = y - 0.5 * v
    return flow(x0, y0)


def rms_error(result, flow, midpoint: bool = True) -> float:
    """RMS of the vector error over valid cells, sqrt(mean(du^2 + dv^2))
    (bench/run_benchmarks.jl:209-220)."""
    X, Y = np.meshgrid(np.asarray(result["x"], float), np.asarray(result["y"], float))
    tu, tv = midpoint_truth(flow, X, Y) if midpoint else flow(X, Y)
    ok = ~np.asarray(result["outliers"], bool)
    if "mask" in result:
        ok &= ~np.asarray(result["mask"], bool)
    e2 = (np.asarray(result["u"], float) - tu) ** 2 + (np.asarray(result["v"], float) - tv) ** 2
    return float(np.sqrt(np.mean(e2[ok])))


def stereo_maps(ny: int = 256, nx: int = 320, image_size=(300, 360)):
    """Two cameras' dewarp maps onto one ny x nx world grid: smooth,
    mildly projective source coordinates (1-based camera row, col per grid
    node), each camera leaving part of the grid out of view (the first
    camera's map runs past its frame's last row, the second's past its last
    column)."""
    i = np.arange(1, ny + 1, dtype=np.float64)[:, None]
    j = np.arange(1, nx + 1, dtype=np.float64)[None, :]
    r1 = 18.0 + 1.08 * i + 0.015 * j + 6e-5 * i * j
    c1 = 12.0 + 0.01 * i + 1.03 * j - 4e-5 * j * j
    r2 = 22.0 + 0.99 * i - 0.02 * j - 5e-5 * i * j
    c2 = 30.0 + 0.012 * i + 1.04 * j + 3e-5 * j * j
    return (np.asfortranarray(r1), np.asfortranarray(c1)), (np.asfortranarray(r2), np.asfortranarray(c2)), image_size


def stereo_frames(image_size=(300, 360), seed: int = 106):
    """Camera frames of a stereo pair: each camera's pair is a particle image
    pair of image_size with its own uniform shift (camera pixels)."""
    nr, nc = image_size
    p1 = particle_pair(nr, nc, uniform_flow(1.6, -0.7), seed, ppp=0.03, noise=0.02)
    p2 = particle_pair(nr, nc, uniform_flow(1.1, -0.9), seed + 1, ppp=0.03, noise=0.02)
    return (p1.A, p1.B), (p2.A, p2.B)
