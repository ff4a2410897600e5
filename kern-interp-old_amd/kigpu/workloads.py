"""Synthetic inputs of the factor/solve workloads (geometry, boundary data,
targets), generated on the host from fixed integer seeds.

The geometries are new Boundary-style inputs, not reference code: Fourier
stars r(t) = 1 + sum_k a_k cos(f_k t + phi_k) sampled at N trapezoidal nodes,
with the reference's conventions for the arrays (``boundaries/boundary.h:
35-51``: interleaved points and normals, outward normal, per-node weight and
signed curvature, positive on a convex outer curve as in ``circle.cpp:21``).
Random draws come from splitmix64 (Steele, Lea and Flood 2014), so the same
seed gives the same inputs in every process.
"""
from __future__ import annotations

import math

import numpy as np

TWO_PI = 2.0 * math.pi
_M64 = (1 << 64) - 1


def uniform01(seed: int, n: int) -> np.ndarray:
    """n doubles in [0, 1): the top 53 bits of successive splitmix64 outputs."""
    out = np.empty(n, dtype=np.float64)
    s = seed & _M64
    for i in range(n):
        s = (s + 0x9E3779B97F4A7C15) & _M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        out[i] = float((z ^ (z >> 31)) >> 11) * (1.0 / (1 << 53))
    return out


def fourier_star(n: int, amps, freqs, phases=None):
    """(points (2n,), normals (2n,), weights (n,), curvatures (n,)) of
    r(t) = 1 + sum_k a_k cos(f_k t + phi_k) at t_j = 2 pi j / n."""
    amps = np.asarray(amps, dtype=np.float64)
    freqs = np.asarray(freqs, dtype=np.float64)
    phases = np.zeros_like(amps) if phases is None else np.asarray(phases, np.float64)
    t = TWO_PI * np.arange(n) / n
    arg = freqs[None, :] * t[:, None] + phases[None, :]
    r = 1.0 + (amps * np.cos(arg)).sum(1)
    dr = (-amps * freqs * np.sin(arg)).sum(1)
    ddr = (-amps * freqs * freqs * np.cos(arg)).sum(1)
    ct, st = np.cos(t), np.sin(t)
    dx = dr * ct - r * st
    dy = dr * st + r * ct
    ddx = ddr * ct - 2 * dr * st - r * ct
    ddy = ddr * st + 2 * dr * ct - r * st
    sp = np.sqrt(dx * dx + dy * dy)
    pts = np.stack([r * ct, r * st], 1).reshape(-1)
    nrm = np.stack([dy / sp, -dx / sp], 1).reshape(-1)
    return pts, nrm, sp * TWO_PI / n, (dx * ddy - ddx * dy) / sp ** 3


def star(n: int):
    """r = 1 + 0.3 cos 5t: the Laplace factor/solve workload."""
    return fourier_star(n, [0.3], [5.0])


def random_star(n: int, seed: int):
    """r = 1 + sum_{k=2..6} a_k cos(k t + phi_k), a_k ~ U(0, 0.05), phi_k ~ U(0, 2 pi)."""
    u = uniform01(seed, 10)
    return fourier_star(n, 0.05 * u[:5], np.arange(2, 7, dtype=np.float64), TWO_PI * u[5:])


def batch_sizes(count: int, seed: int, lo: int = 1024, hi: int = 16384, step: int = 64):
    """count sizes ~ U{lo..hi} rounded down to a multiple of step."""
    u = uniform01(seed, count)
    n = lo + np.floor(u * (hi - lo + 1)).astype(np.int64)
    return np.maximum(lo, (n // step) * step)


def point_source_bc(points, src=(3.0, 3.0)):
    """f = log|x - src| / (2 pi) at the nodes."""
    xy = np.asarray(points).reshape(-1, 2)
    return np.log(np.sqrt((xy[:, 0] - src[0]) ** 2 + (xy[:, 1] - src[1]) ** 2)) / TWO_PI


def point_sources(count: int, seed: int, rmin=2.0, rmax=4.0) -> np.ndarray:
    """count source locations, radius ~ U(rmin, rmax), angle ~ U(0, 2 pi)."""
    u = uniform01(seed, 2 * count).reshape(count, 2)
    rho = rmin + (rmax - rmin) * u[:, 0]
    phi = TWO_PI * u[:, 1]
    return np.stack([rho * np.cos(phi), rho * np.sin(phi)], 1)


def interior_disc_targets(n: int, seed: int, radius: float = 0.6) -> np.ndarray:
    """n targets uniform in the disc |x| < radius, interleaved (2n,)."""
    u = uniform01(seed, 2 * n).reshape(n, 2)
    rho = radius * np.sqrt(u[:, 0])
    phi = TWO_PI * u[:, 1]
    return np.stack([rho * np.cos(phi), rho * np.sin(phi)], 1).reshape(-1)
