"""Pins of oracle/cubic.py (the spline boundaries and the in-domain rules).

* Nodes, normals and curvatures against scipy's periodic CubicSpline through
  the same knots at uniform parameter: the periodic C2 cubic interpolant is
  unique, so an independent solver must give the same curve.  A dropped term
  of the cubic, a wrong sign of the right-hand side or a shifted knot fails.
* Weights: their sum against the arc length of that spline (the polyline of
  10 samples per node is within 1e-5 relative of it at these sizes), and each
  weight against the mean arc length of the two pieces next to its node.
* The crossing rule against an exact even-odd polygon test (intersection
  abscissae of every edge) on points further than 0.03 from every node,
  where the reference's vertex-based count and the exact one agree.
* The disc rules at points placed either side of the eps band.
"""
import math

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.interpolate import CubicSpline

from oracle import cubic as C
from oracle import geometry as G


def spline_of(xs, ys):
    n = len(xs)
    u = np.arange(n + 1, dtype=np.float64)
    pts = np.stack([np.r_[xs, xs[0]], np.r_[ys, ys[0]]], 1)
    return CubicSpline(u, pts, bc_type="periodic")


def starfish_knots(scale=0.375, c=(0.5, 0.5)):
    return C._star_knots(c[0], c[1], scale)


@pytest.mark.parametrize("N", [1024, 8192])
def test_starfish_against_scipy_spline(N):
    b = C.starfish(N)
    nps = N // 20
    assert b.num_points == 20 * nps == b.num_outer
    sp = spline_of(*starfish_knots())
    u = np.array([i + j / nps for i in range(20) for j in range(nps)])
    p, d1, d2 = sp(u), sp(u, 1), sp(u, 2)
    assert np.abs(b.xy() - p).max() < 1e-13
    speed = np.hypot(d1[:, 0], d1[:, 1])
    nrm = np.stack([d1[:, 1], -d1[:, 0]], 1) / speed[:, None]
    assert np.abs(b.normals.reshape(-1, 2) - nrm).max() < 1e-12
    kap = (d1[:, 0] * d2[:, 1] - d2[:, 0] * d1[:, 1]) / speed ** 3
    assert np.abs(b.curvatures - kap).max() < 1e-10 * np.abs(kap).max()
    # outward normals of a counter-clockwise curve: Gauss-Green area > 0
    xy = b.xy()
    area = 0.5 * np.sum(xy[:, 0] * nrm[:, 0] * b.weights + xy[:, 1] * nrm[:, 1] * b.weights)
    assert area > 0


def test_starfish_weights_against_arc_length():
    N = 1024
    b = C.starfish(N)
    nps = N // 20
    sp = spline_of(*starfish_knots())
    speed = lambda u: float(np.hypot(*sp(u, 1)))
    total = sum(quad(speed, i, i + 1, limit=200)[0] for i in range(20))
    assert abs(b.weights.sum() - total) <= 1e-5 * total
    # weight i = mean of the arc lengths of the pieces before and after node i
    h = 1.0 / nps
    for i in [0, 1, 7, nps - 1, nps, 5 * nps + 3, 20 * nps - 1]:
        u0 = i * h
        after = quad(speed, u0, u0 + h)[0]
        before = quad(speed, (u0 - h) % 20, (u0 - h) % 20 + h)[0]
        assert abs(b.weights[i] - 0.5 * (before + after)) <= 1e-4 * b.weights[i]


@pytest.mark.parametrize("name", ["ex1", "ex2", "ex3"])
def test_experiment_domains_sections(name):
    b = getattr(C, name)(4096)
    assert b.num_outer + sum(h.num_nodes for h in b.holes) == b.num_points
    o = b.num_outer
    for h in b.holes:
        xy = b.xy()[o:o + h.num_nodes]
        nm = b.normals.reshape(-1, 2)[o:o + h.num_nodes]
        # hole normals point into the hole: towards its centre
        to_c = np.array(h.center)[None, :] - xy
        assert np.all(np.sum(nm * to_c, 1) > 0)
        if name != "ex3" and np.allclose(np.hypot(*(xy - h.center).T), h.radius):
            # circular holes: exact weights and the closed-form perimeter
            assert abs(b.weights[o:o + h.num_nodes].sum() - 2 * math.pi * h.radius) < 1e-13
        o += h.num_nodes
    # outer normals point away from the hole centres' convex hull side: the
    # outer curve is counter-clockwise (positive signed area)
    xy = b.xy()[:b.num_outer]
    area = 0.5 * np.sum(xy[:, 0] * np.roll(xy[:, 1], -1) - np.roll(xy[:, 0], -1) * xy[:, 1])
    assert area > 0


def test_experiment_boundary_data():
    b = C.ex3(4096, "EX3A")
    o, h1, h2 = b.num_outer, b.holes[0].num_nodes, b.holes[1].num_nodes
    f = b.boundary_values
    assert f.shape == (b.num_points,)
    assert np.all(f[:o] == 0) and np.all(f[o:o + h1] == 1) and np.all(f[o + h1:] == -1)
    b = C.ex3(4096, "EX3B")
    f = b.boundary_values.reshape(-1, 2)
    nm = b.normals.reshape(-1, 2)
    assert np.all(f[:o] == [1.0, 0.0])
    assert np.array_equal(f[o:o + h1], -nm[o:o + h1]) and np.array_equal(f[o + h1:], nm[o + h1:])
    b = C.ex2(4096)
    f = b.boundary_values.reshape(-1, 2)
    x = b.xy()[:, 0]
    assert np.array_equal(f[:, 0], ((x < -0.99) | (x > 1.99)).astype(float))
    assert f[:, 0].sum() > 0 and np.all(f[:, 1] == 0)


def even_odd(poly, t):
    """Exact crossing-number test of points t against a closed polygon."""
    x, y = poly[:, 0], poly[:, 1]
    xn, yn = np.roll(x, -1), np.roll(y, -1)
    inside = np.zeros(t.shape[0], bool)
    for a, b_, c, d in zip(x, y, xn, yn):
        straddle = (b_ > t[:, 1]) != (d > t[:, 1])
        with np.errstate(divide="ignore", invalid="ignore"):
            xi = a + (t[:, 1] - b_) * (c - a) / (d - b_)
        inside ^= straddle & (xi > t[:, 0])
    return inside


def far_from_nodes(b, t, dmin):
    xy = b.xy()
    keep = np.ones(t.shape[0], bool)
    for s in range(0, xy.shape[0], 512):
        d = np.sqrt(((t[:, None, :] - xy[None, s:s + 512, :]) ** 2).sum(-1)).min(1)
        keep &= d > dmin
    return t[keep]


@pytest.mark.parametrize("name", ["starfish", "ex1", "ex3"])
def test_winding_rule_against_exact_polygon_test(name):
    N = 2048
    b = getattr(C, name)(N)
    u = G.uniform01(17, 8000).reshape(-1, 2)
    lo, hi = b.xy().min(0) - 0.05, b.xy().max(0) + 0.05
    t = far_from_nodes(b, lo + u * (hi - lo), 0.03)
    assert t.shape[0] > 1000
    got = C.in_winding_domain(b, t)
    want = even_odd(b.xy()[:b.num_outer], t)
    o = b.num_outer
    for h in b.holes:
        want &= ~even_odd(b.xy()[o:o + h.num_nodes], t)
        o += h.num_nodes
    assert np.array_equal(got, want)
    assert 0.1 < got.mean() < 0.9
    # eps: a point next to a node is rejected
    assert not C.in_winding_domain(b, b.xy()[5:6] + 0.005)[0]


def test_winding_rule_remainder_quirk():
    # C's truncating %: a hole traversed clockwise counts -1 there and does
    # not exclude (cubic_boundary.cpp:222 tests == 1)
    assert C._cmod(-1, 2) == -1 and C._cmod(-2, 2) == 0 and C._cmod(3, 2) == 1
    b = C.starfish(1024)
    ang = np.linspace(0, 2 * np.pi, 64, endpoint=False)
    ring = np.stack([0.5 + 0.1 * np.cos(ang), 0.5 + 0.1 * np.sin(ang)], 1)
    for order, excluded in [(1, True), (-1, False)]:
        hb = G.Boundary(np.r_[b.points, ring[::order].reshape(-1)], None,
                        np.r_[b.weights, np.ones(64)], None,
                        holes=[G.Hole((0.5, 0.5), 0.1, 64)])
        hb.num_outer = b.num_outer
        assert C.in_winding_domain(hb, np.array([[0.52, 0.513]]))[0] == (not excluded)


def test_disc_rules_closed_forms():
    eps = 1e-2
    circ = [(0.5, 0.5 + 0.25 - eps - 1e-6, True), (0.5, 0.5 + 0.25 - eps + 1e-6, False),
            (0.5 + 0.1, 0.5, True), (0.9, 0.9, False)]
    for x, y, want in circ:
        assert C.in_disc_domain([x, y], (0.5, 0.5), 0.25)[0] == want
    ann = G.annulus(1000)
    # hole at (0.9, 0.5) radius 0.2: excluded within 0.2 + eps of its centre
    for x, y, want in [(0.9, 0.5 + 0.2 + eps + 1e-6, True), (0.9, 0.5 + 0.2 + eps - 1e-6, False),
                       (0.5, 1.5 - eps - 1e-6, True), (0.5, 1.5 - eps + 1e-6, False)]:
        assert C.in_disc_domain([x, y], (0.5, 0.5), 1.0, ann.holes)[0] == want
