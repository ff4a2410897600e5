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


def circle_ring(center, radius, n, into_hole):
    """(points, normals, weights, curvatures) of n equispaced nodes on a circle,
    the conventions of the reference's circle and annulus boundaries
    (circle.cpp:11-22, annulus.cpp:37-52): outward normals and curvature 1/R on
    the outer curve; on a hole the normal points into the hole (out of the
    domain) and the curvature is -1/R."""
    ang = np.array([i * TWO_PI / n for i in range(n)])
    c, s = np.cos(ang), np.sin(ang)
    pts = np.stack([center[0] + radius * c, center[1] + radius * s], 1).reshape(-1)
    sg = -1.0 if into_hole else 1.0
    nrm = np.stack([sg * c, sg * s], 1).reshape(-1)
    w = np.full(n, 2.0 * radius * math.pi / n)
    kap = np.full(n, -1.0 / radius if into_hole else 1.0 / radius)
    return pts, nrm, w, kap


def holed_disc(n_outer: int, n_hole: int, n_holes: int, seed: int):
    """The unit disc at the origin with n_holes seeded circular holes, radii
    ~ U(0.05, 0.15), centres by rejection sampling with a gap of 0.05 to each
    other and to the outer circle.  Returns ((points, normals, weights,
    curvatures), holes) with holes a list of (cx, cy, radius, n_hole)."""
    u = uniform01(seed, 4096)
    holes = []
    j = 0
    while len(holes) < n_holes:
        rad = 0.05 + 0.10 * u[j]
        cx = -1.0 + 2.0 * u[j + 1]
        cy = -1.0 + 2.0 * u[j + 2]
        j += 3
        if math.hypot(cx, cy) + rad > 1.0 - 0.05:
            continue
        if any(math.hypot(cx - h[0], cy - h[1]) < rad + h[2] + 0.05 for h in holes):
            continue
        holes.append((cx, cy, rad, n_hole))
    parts = [circle_ring((0.0, 0.0), 1.0, n_outer, False)]
    parts += [circle_ring((h[0], h[1]), h[2], h[3], True) for h in holes]
    geom = tuple(np.concatenate([p[q] for p in parts]) for q in range(4))
    return geom, holes


def border_matrices(stokes: bool, holes, points, weights):
    """QuadTree::U and ::Psi of a multiply connected domain
    (initialize_U_mat, linear_solve.cpp:18-64; initialize_Psi_mat, :67-110).
    Laplace: U(i, h) = log|x_i - c_h|, Psi(h, i) = w_i on hole h's nodes.
    Stokes (2 DOFs per node, 3 unknowns per hole): the Stokeslet and rotlet
    columns of U, and Psi rows (w, 0), (0, w), (w y, -w x).  A node belongs to
    the first hole with |x - c| < radius + 1e-8."""
    xy = np.asarray(points, np.float64).reshape(-1, 2)
    w = np.asarray(weights, np.float64)
    npt = xy.shape[0]
    owner = np.full(npt, -1)
    for h in range(len(holes) - 1, -1, -1):
        d = np.sqrt((xy[:, 0] - holes[h][0]) ** 2 + (xy[:, 1] - holes[h][1]) ** 2)
        owner[d < holes[h][2] + 1e-8] = h
    if not stokes:
        U = np.empty((npt, len(holes)))
        for h, hole in enumerate(holes):
            U[:, h] = np.log(np.sqrt((xy[:, 0] - hole[0]) ** 2 + (xy[:, 1] - hole[1]) ** 2))
        Psi = np.zeros((len(holes), npt))
        idx = np.nonzero(owner >= 0)[0]
        Psi[owner[idx], idx] = w[idx]
        return U, Psi
    U = np.zeros((2 * npt, 3 * len(holes)))
    sc = 1.0 / (4 * math.pi)
    for h, hole in enumerate(holes):
        rx, ry = xy[:, 0] - hole[0], xy[:, 1] - hole[1]
        nr = np.sqrt(rx * rx + ry * ry)
        inv2 = 1.0 / nr ** 2
        U[0::2, 3 * h] = sc * (np.log(1 / nr) + inv2 * rx * rx)
        U[1::2, 3 * h] = sc * (inv2 * ry * rx)
        U[0::2, 3 * h + 1] = sc * (inv2 * rx * ry)
        U[1::2, 3 * h + 1] = sc * (np.log(1 / nr) + inv2 * ry * ry)
        U[0::2, 3 * h + 2] = ry * (sc / nr ** 2)
        U[1::2, 3 * h + 2] = -rx * (sc / nr ** 2)
    Psi = np.zeros((3 * len(holes), 2 * npt))
    for i in np.nonzero(owner >= 0)[0]:
        h = owner[i]
        Psi[3 * h, 2 * i] = w[i]
        Psi[3 * h + 1, 2 * i + 1] = w[i]
        Psi[3 * h + 2, 2 * i] = w[i] * xy[i, 1]
        Psi[3 * h + 2, 2 * i + 1] = -w[i] * xy[i, 0]
    return U, Psi


def stokeslet_bc(points, src=(2.0, 1.0), force=(1.0, 0.0)):
    """Velocity (interleaved u_x, u_y) at the nodes of a Stokeslet at src with
    the given force: u = (1 / 4 pi) (-log r I + r r^T / r^2) F."""
    xy = np.asarray(points).reshape(-1, 2)
    rx, ry = xy[:, 0] - src[0], xy[:, 1] - src[1]
    r2 = rx * rx + ry * ry
    lg = -0.5 * np.log(r2)
    fx, fy = force
    ux = (lg * fx + (rx * rx * fx + rx * ry * fy) / r2) / (4 * math.pi)
    uy = (lg * fy + (ry * rx * fx + ry * ry * fy) / r2) / (4 * math.pi)
    return np.stack([ux, uy], 1).reshape(-1)
