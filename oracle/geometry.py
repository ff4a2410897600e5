"""Boundary geometries and boundary conditions used as test inputs.

TEST INFRASTRUCTURE ONLY (see the package docstring).  The product path takes
these arrays as plain inputs; this module only generates them.

A boundary is the reference's ``Boundary`` input contract
(``kern_interp/boundaries/boundary.h:35-51``): interleaved points and normals
``[x0, y0, x1, y1, ...]``, per-node quadrature weights and curvatures, the
list of circular holes and the right-hand side ``boundary_values``.

Pinned by: the circle closed form (every off-diagonal Laplace double-layer
entry on a circle equals ``w / (4 pi R)``), the quadrature identity
``sum(w) = perimeter`` and Gauss' law for the double layer
(``tests/test_oracle_kernel.py``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

TWO_PI = 2.0 * math.pi


@dataclass
class Hole:
    """Circular hole: ``kern_interp/boundaries/boundary.h:12-16``."""
    center: tuple
    radius: float
    num_nodes: int = 0


@dataclass
class Boundary:
    points: np.ndarray          # (2N,) interleaved x, y
    normals: np.ndarray         # (2N,) interleaved, outward from the domain
    weights: np.ndarray         # (N,)
    curvatures: np.ndarray      # (N,)
    holes: list = field(default_factory=list)
    boundary_values: np.ndarray | None = None   # (sd*N,)
    name: str = ""

    @property
    def num_points(self) -> int:
        return self.weights.shape[0]

    def xy(self) -> np.ndarray:
        return self.points.reshape(-1, 2)


# --------------------------------------------------------------------------
# counter-based RNG shared by tests and bench (splitmix64)
# --------------------------------------------------------------------------
_M64 = (1 << 64) - 1


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n uint64 draws of splitmix64 starting from ``seed`` (Steele et al. 2014)."""
    out = np.empty(n, dtype=np.uint64)
    s = seed & _M64
    for i in range(n):
        s = (s + 0x9E3779B97F4A7C15) & _M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        out[i] = z ^ (z >> 31)
    return out


def uniform01(seed: int, n: int) -> np.ndarray:
    """n doubles in [0, 1) from the top 53 bits of splitmix64."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


# --------------------------------------------------------------------------
# reference test geometries (restated from kern_interp/boundaries/*.cpp)
# --------------------------------------------------------------------------
def circle(n: int) -> Boundary:
    """Circle of radius 0.25 centred at (0.5, 0.5): ``circle.cpp:8-23``."""
    ang = np.array([i * 2.0 * math.pi / n for i in range(n)])
    c, s = np.cos(ang), np.sin(ang)
    pts = np.stack([0.5 + 0.25 * c, 0.5 + 0.25 * s], axis=1).reshape(-1)
    nrm = np.stack([c, s], axis=1).reshape(-1)
    return Boundary(pts, nrm, np.full(n, math.pi / (n * 2)), np.full(n, 4.0),
                    name="circle")


def _ring(center, radius, n, inward):
    ang = np.array([i * 2.0 * math.pi / n for i in range(n)])
    c, s = np.cos(ang), np.sin(ang)
    pts = np.stack([center[0] + radius * c, center[1] + radius * s], 1).reshape(-1)
    sgn = -1.0 if inward else 1.0
    nrm = np.stack([sgn * c, sgn * s], 1).reshape(-1)
    w = np.full(n, 2 * radius * math.pi / n)
    k = np.full(n, (-1.0 / radius) if inward else (1.0 / radius))
    return pts, nrm, w, k


def _with_holes(n_outer, outer_c, outer_r, holes, name):
    pts, nrm, w, k = _ring(outer_c, outer_r, n_outer, inward=False)
    parts = [(pts, nrm, w, k)]
    for h in holes:
        parts.append(_ring(h.center, h.radius, h.num_nodes, inward=True))
    return Boundary(np.concatenate([p[0] for p in parts]),
                    np.concatenate([p[1] for p in parts]),
                    np.concatenate([p[2] for p in parts]),
                    np.concatenate([p[3] for p in parts]),
                    holes=list(holes), name=name)


def annulus(n: int) -> Boundary:
    """Unit circle at (0.5, 0.5) with two holes of radius 0.2: ``annulus.cpp:8-63``."""
    hn = n // 5
    holes = [Hole((0.9, 0.5), 0.2, hn), Hole((0.1, 0.5), 0.2, hn)]
    return _with_holes(n, (0.5, 0.5), 1.0, holes, "annulus")


def donut(n: int) -> Boundary:
    """Unit circle at (0.5, 0.5) with one concentric hole of radius 0.5: ``donut.cpp:8-64``."""
    hn = n // 5
    return _with_holes(n, (0.5, 0.5), 1.0, [Hole((0.5, 0.5), 0.5, hn)], "donut")


# --------------------------------------------------------------------------
# workload geometries (no reference counterpart; new Boundary subclasses)
# --------------------------------------------------------------------------
def fourier_star(n: int, amps, freqs, phases=None, name="star") -> Boundary:
    """Trapezoidal nodes on r(t) = 1 + sum_k a_k cos(f_k t + phi_k).

    Outward normal n = (tau_y, -tau_x)/|tau| for the counter-clockwise
    tangent tau, weight |tau| 2pi/N, signed curvature
    (x'y'' - x''y')/|tau|^3 (positive on a convex outer curve, matching the
    reference's circle, ``circle.cpp:21``).
    """
    amps = np.asarray(amps, dtype=np.float64)
    freqs = np.asarray(freqs, dtype=np.float64)
    phases = np.zeros_like(amps) if phases is None else np.asarray(phases, np.float64)
    t = TWO_PI * np.arange(n) / n
    arg = freqs[None, :] * t[:, None] + phases[None, :]
    r = 1.0 + (amps * np.cos(arg)).sum(1)
    dr = (-amps * freqs * np.sin(arg)).sum(1)
    ddr = (-amps * freqs * freqs * np.cos(arg)).sum(1)
    ct, st = np.cos(t), np.sin(t)
    x, y = r * ct, r * st
    dx = dr * ct - r * st
    dy = dr * st + r * ct
    ddx = ddr * ct - 2 * dr * st - r * ct
    ddy = ddr * st + 2 * dr * ct - r * st
    sp = np.sqrt(dx * dx + dy * dy)
    pts = np.stack([x, y], 1).reshape(-1)
    nrm = np.stack([dy / sp, -dx / sp], 1).reshape(-1)
    w = sp * TWO_PI / n
    kap = (dx * ddy - ddx * dy) / sp ** 3
    return Boundary(pts, nrm, w, kap, name=name)


def star(n: int) -> Boundary:
    """The star r = 1 + 0.3 cos(5 theta) (Laplace factor/solve workload)."""
    return fourier_star(n, [0.3], [5.0], name="star")


def random_star(n: int, seed: int) -> Boundary:
    """r = 1 + sum_{k=2..6} a_k cos(k t + phi_k), a_k ~ U(0, 0.05), phi_k ~ U(0, 2pi)."""
    u = uniform01(seed, 10)
    amps = 0.05 * u[:5]
    phases = TWO_PI * u[5:]
    return fourier_star(n, amps, np.arange(2, 7, dtype=np.float64), phases,
                        name=f"rstar{seed}")


def multi_hole_disc(n_outer: int, n_hole: int, n_holes: int, seed: int) -> Boundary:
    """Unit disc at the origin with ``n_holes`` seeded non-overlapping circular holes.

    Radii ~ U(0.05, 0.15); centres by rejection sampling so that holes keep a
    gap of 0.05 from each other and from the outer circle.  Hole normals point
    into the hole (out of the domain), as in ``annulus.cpp:47-48``.
    """
    holes = []
    u = uniform01(seed, 4096)
    j = 0
    while len(holes) < n_holes:
        rad = 0.05 + 0.10 * u[j]
        cx = -1.0 + 2.0 * u[j + 1]
        cy = -1.0 + 2.0 * u[j + 2]
        j += 3
        if math.hypot(cx, cy) + rad > 1.0 - 0.05:
            continue
        if any(math.hypot(cx - h.center[0], cy - h.center[1]) < rad + h.radius + 0.05
               for h in holes):
            continue
        holes.append(Hole((cx, cy), rad, n_hole))
    return _with_holes(n_outer, (0.0, 0.0), 1.0, holes, f"holes{seed}")


# --------------------------------------------------------------------------
# boundary conditions (restated from kern_interp/boundaries/boundary.cpp:43-120)
# --------------------------------------------------------------------------
def bc_point_source(b: Boundary, src=(3.0, 3.0)) -> np.ndarray:
    """f = log|x - src| / (2 pi); ``SINGLE_ELECTRON`` (boundary.cpp:49-54) is src = (-3, -2)."""
    xy = b.xy()
    return np.log(np.sqrt((xy[:, 0] - src[0]) ** 2 + (xy[:, 1] - src[1]) ** 2)) / TWO_PI


def bc_single_electron(b: Boundary) -> np.ndarray:
    return bc_point_source(b, (-3.0, -2.0))


def bc_tangent(b: Boundary) -> np.ndarray:
    """TANGENT_VEC: (-n_y, n_x) per node (boundary.cpp:70-73)."""
    n = b.normals.reshape(-1, 2)
    return np.stack([-n[:, 1], n[:, 0]], 1).reshape(-1)


def bc_stokeslet(b: Boundary, src=(2.0, 1.0), force=(1.0, 0.0)) -> np.ndarray:
    """Velocity of a Stokeslet at ``src`` with ``force``: u = (1/4pi)(-log r I + r r^T / r^2) F."""
    xy = b.xy()
    rx = xy[:, 0] - src[0]
    ry = xy[:, 1] - src[1]
    r2 = rx * rx + ry * ry
    lg = -0.5 * np.log(r2)
    fx, fy = force
    ux = (lg * fx + (rx * rx * fx + rx * ry * fy) / r2) / (4 * math.pi)
    uy = (lg * fy + (ry * rx * fx + ry * ry * fy) / r2) / (4 * math.pi)
    return np.stack([ux, uy], 1).reshape(-1)


def interior_disc_targets(n: int, seed: int, radius: float = 0.6) -> np.ndarray:
    """n targets uniform in the disc |x| < radius: rho = radius sqrt(u1), phi = 2 pi u2."""
    u = uniform01(seed, 2 * n).reshape(n, 2)
    rho = radius * np.sqrt(u[:, 0])
    phi = TWO_PI * u[:, 1]
    return np.stack([rho * np.cos(phi), rho * np.sin(phi)], 1).reshape(-1)


def point_sources(nrhs: int, seed: int, rmin=2.0, rmax=4.0) -> np.ndarray:
    """nrhs source locations with radius ~ U(rmin, rmax), angle ~ U(0, 2 pi)."""
    u = uniform01(seed, 2 * nrhs).reshape(nrhs, 2)
    rho = rmin + (rmax - rmin) * u[:, 0]
    phi = TWO_PI * u[:, 1]
    return np.stack([rho * np.cos(phi), rho * np.sin(phi)], 1)
