"""Dense shape-gradient oracle for the Laplace double layer.

TEST INFRASTRUCTURE ONLY (see the package docstring).

The reference has no adjoint: its experiments differentiate by finite
differences of whole solves (``experiments/experiment_three.cpp:22-56,
160-171``).  This module computes the same gradients densely and by
complex-step differentiation, so it shares no formula with the adjoint code
under test:

* the curve r(t) = 1 + 0.3 cos 5t + sum_k (c_{2k} cos (k+1)t + c_{2k+1} sin (k+1)t)
  is discretised like :func:`oracle.geometry.fourier_star` (trapezoidal nodes,
  outward normal, weight |tau| 2pi/N, curvature (x'y'' - x''y')/|tau|^3), in
  arithmetic that accepts complex coefficients;
* K is the dense Laplace double-layer matrix of ``kernel.cpp:11-36``, the
  forward matrix F the same entries at domain targets (``kernel.cpp:337-357``),
  f = x^2 + x y on the boundary (not harmonic, so the interior field and the
  loss depend on the shape; data from a harmonic function such as a point
  source would give u independent of the curve and a zero gradient);
* L(c) = sum_j u_j^2 with u = F K^{-1} f, and dL/dc_k = Im L(c + i h e_k) / h
  (complex step, h = 1e-30: no cancellation, exact to rounding).

The complex step itself is pinned against central finite differences of the
real loss (``tests/test_oracle_adjoint.py``).
"""
from __future__ import annotations

import numpy as np

TWO_PI = 2.0 * np.pi
INV_2PI = 1.0 / TWO_PI


def param_star(n: int, c) -> dict:
    """Geometry of the parameterised star (SoA, complex when c is complex)."""
    c = np.asarray(c)
    t = TWO_PI * np.arange(n) / n
    r = 1.0 + 0.3 * np.cos(5 * t) + 0j * c.sum()
    dr = -1.5 * np.sin(5 * t) + 0j * c.sum()
    ddr = -7.5 * np.cos(5 * t) + 0j * c.sum()
    for k in range(len(c) // 2):
        f = k + 1.0
        a, b = c[2 * k], c[2 * k + 1]
        r = r + a * np.cos(f * t) + b * np.sin(f * t)
        dr = dr - a * f * np.sin(f * t) + b * f * np.cos(f * t)
        ddr = ddr - a * f * f * np.cos(f * t) - b * f * f * np.sin(f * t)
    ct, st = np.cos(t), np.sin(t)
    x, y = r * ct, r * st
    dx = dr * ct - r * st
    dy = dr * st + r * ct
    ddx = ddr * ct - 2 * dr * st - r * ct
    ddy = ddr * st + 2 * dr * ct - r * st
    sp = np.sqrt(dx * dx + dy * dy)
    g = {"x": x, "y": y, "nx": dy / sp, "ny": -dx / sp, "w": sp * TWO_PI / n,
         "kappa": (dx * ddy - ddx * dy) / sp ** 3}
    if not np.iscomplexobj(c) or np.all(np.imag(c) == 0):
        g = {k: np.real(v) for k, v in g.items()}
    return g


def dense_K(g) -> np.ndarray:
    """K(i, j) of ``Kernel::laplace`` (kernel.cpp:11-36), targets = sources = the nodes."""
    r1 = g["x"][:, None] - g["x"][None, :]
    r2 = g["y"][:, None] - g["y"][None, :]
    return _dl(r1, r2, g)


def dense_F(g, targets) -> np.ndarray:
    """Forward matrix at domain targets (kernel.cpp:337-357)."""
    xy = np.asarray(targets, dtype=np.float64).reshape(-1, 2)
    r1 = xy[:, 0][:, None] - g["x"][None, :]
    r2 = xy[:, 1][:, None] - g["y"][None, :]
    return _dl(r1, r2, g)


def _dl(r1, r2, g):
    diag = (r1 == 0) & (r2 == 0)
    rr = np.where(diag, 1.0, r1 * r1 + r2 * r2)
    off = -g["w"][None, :] * INV_2PI * (r1 * g["nx"][None, :] + r2 * g["ny"][None, :]) / rr
    dval = 0.5 + 0.5 * g["kappa"] * g["w"] * INV_2PI
    return np.where(diag, np.broadcast_to(dval[None, :], off.shape), off)


def quadratic_bc(g):
    return g["x"] * g["x"] + g["x"] * g["y"]


def dense_loss(n, c, targets):
    """L(c) = sum_j u_j^2, u = F K^{-1} f (dense LAPACK solve)."""
    g = param_star(n, c)
    mu = np.linalg.solve(dense_K(g), quadratic_bc(g))
    u = dense_F(g, targets) @ mu
    return np.sum(u * u)


def complex_step_grad(n, c, targets, h=1e-30):
    c = np.asarray(c, dtype=np.float64)
    out = np.zeros(c.size)
    for k in range(c.size):
        cc = c.astype(np.complex128)
        cc[k] += 1j * h
        out[k] = np.imag(dense_loss(n, cc, targets)) / h
    return out


def fd_grad(n, c, targets, h=1e-6):
    c = np.asarray(c, dtype=np.float64)
    out = np.zeros(c.size)
    for k in range(c.size):
        e = np.zeros(c.size)
        e[k] = h
        out[k] = (dense_loss(n, c + e, targets) - dense_loss(n, c - e, targets)) / (2 * h)
    return out


def geometry_directional(g, dg, fn, h=1e-30):
    """d/ds fn(g + s dg) at s = 0 by complex step; g, dg: dicts of SoA arrays."""
    gc = {k: g[k] + 1j * h * dg[k] for k in g}
    return np.imag(fn(gc)) / h
