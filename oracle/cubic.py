"""Cubic-spline boundaries and the in-domain rules of the reference's geometries.

TEST INFRASTRUCTURE ONLY (see the package docstring): the product path gets
boundary arrays as inputs and runs the in-domain rules on the GPU; this module
generates the arrays and states the rules the GPU results are checked against.

* ``CubicBoundary`` (kern_interp/boundaries/cubic_boundary.cpp:10-178): a
  closed curve through spline knots, one periodic C2 cubic per knot interval
  (the cyclic system 4 d_i + d_{i-1} + d_{i+1} = 3 (p_{i+1} - p_{i-1}),
  :10-77), nodes at t = j / nodes_per_spline of every interval, unit normals
  (y', -x') / |p'| (negated for a hole), signed curvature
  (x'y'' - x''y') / |p'|^3 (negated for a hole) and weights from a polyline of
  10 samples per node (:90-178).
* The starfish ``CubicSpline`` (cubic_spline.cpp:13-50) and the experiment
  domains ``Ex1Boundary`` (ex1boundary.cpp:10-173), ``Ex2Boundary``
  (ex2boundary.cpp:10-94), ``Ex3Boundary`` (ex3boundary.cpp:10-132), with
  their boundary data (boundary.cpp:43-120, applied per node range).
* ``is_in_domain`` of Circle (circle.cpp:34-42), Annulus (annulus.cpp:66-78),
  Donut (donut.cpp:67-79) and CubicBoundary (cubic_boundary.cpp:182-230: the
  crossing count of the node polygon, kept as the reference writes it: only
  nodes to the right of the target count, the outer section must be odd, a
  hole section excludes the target when its count is exactly +1 under C's
  truncating remainder, a node within eps = 1e-2 excludes it).

Pinned by tests/test_oracle_cubic.py: scipy's periodic CubicSpline through the
same knots (an independent solver of the same C2 interpolant) for the nodes,
normals and curvatures; the arc length of that spline for the weight sum; an
exact even-odd polygon test on points away from the curve for the crossing
rule; the closed forms of the circle rules.
"""
from __future__ import annotations

import math

import numpy as np

from oracle.geometry import Boundary, Hole

NUM_SPLINE_POINTS = 20    # cubic_spline.cpp:8
QUAD_SIZE = 10            # cubic_boundary.cpp:102
EPS_DOMAIN = 1e-2         # every is_in_domain rule of the reference


# --------------------------------------------------------------------------
# CubicBoundary
# --------------------------------------------------------------------------
def get_cubics(xk, yk):
    """Per-interval cubic coefficients [c0, c1, c2, c3] in x and y
    (cubic_boundary.cpp:10-77)."""
    n = len(xk)
    A = np.zeros((n, n))
    for i in range(n):
        A[i, i] = 4.0
        if i:
            A[i, i - 1] = 1.0
        if i != n - 1:
            A[i, i + 1] = 1.0
    A[0, n - 1] = 1.0
    A[n - 1, 0] = 1.0
    out = []
    for p in (list(xk), list(yk)):
        rhs = np.array([3 * (p[(i + 1) % n] - p[(i + n - 1) % n]) for i in range(n)])
        d = np.linalg.solve(A, rhs)      # dgesv: the LU solve of left_multiply_inverse
        cub = []
        for i in range(n):
            j = (i + 1) % n
            cub.append([p[i], d[i], 3 * (p[j] - p[i]) - 2 * d[i] - d[j],
                        2 * (p[i] - p[j]) + d[i] + d[j]])
        out.append(cub)
    return out[0], out[1]


def _cubic(c, t):
    return c[0] + t * c[1] + math.pow(t, 2) * c[2] + math.pow(t, 3) * c[3]


def interpolate(b: dict, is_interior: bool, nps: int, xc, yc):
    """Append the nodes of one closed spline curve to ``b`` (cubic_boundary.cpp:90-178)."""
    nsp = len(xc)
    quad = []
    for i in range(nsp):
        cx, cy = xc[i], yc[i]
        for j in range(nps):
            t = j / (nps + 0.0)
            b["p"] += [_cubic(cx, t), _cubic(cy, t)]
            for q in range(QUAD_SIZE):
                tq = ((j * QUAD_SIZE) + q) / (0.0 + QUAD_SIZE * nps)
                quad += [_cubic(cx, tq), _cubic(cy, tq)]
            xp = cx[1] + 2 * t * cx[2] + 3 * math.pow(t, 2) * cx[3]
            yp = cy[1] + 2 * t * cy[2] + 3 * math.pow(t, 2) * cy[3]
            xpp = 2 * cx[2] + 6 * t * cx[3]
            ypp = 2 * cy[2] + 6 * t * cy[3]
            sp = math.sqrt(math.pow(xp, 2) + math.pow(yp, 2))
            kap = (xp * ypp - xpp * yp) / math.pow(sp, 3)
            xp /= sp
            yp /= sp
            if is_interior:
                b["k"].append(-kap)
                b["n"] += [-yp, xp]
            else:
                b["k"].append(kap)
                b["n"] += [yp, -xp]
    nn = nsp * nps
    L = len(quad)
    dist = []
    for i in range(nn):
        s = 0.0
        for j in range(QUAD_SIZE):
            idx = 2 * (QUAD_SIZE * i + j)
            s += math.sqrt(math.pow(quad[idx] - quad[(idx + 2) % L], 2)
                           + math.pow(quad[idx + 1] - quad[(idx + 3) % L], 2))
        dist.append(s)
    b["w"].append((dist[-1] + dist[0]) / 2.0)
    for i in range(1, nn):
        b["w"].append((dist[i - 1] + dist[i]) / 2.0)


def _new():
    return {"p": [], "n": [], "w": [], "k": []}


def _finish(b: dict, holes, num_outer, name) -> Boundary:
    out = Boundary(np.array(b["p"]), np.array(b["n"]), np.array(b["w"]), np.array(b["k"]),
                   holes=list(holes), name=name)
    out.num_outer = num_outer
    return out


def _circle_hole(b: dict, h: Hole):
    """A circular hole of the experiment domains (ex1boundary.cpp:128-147)."""
    n = h.num_nodes
    for i in range(n):
        ang = (2.0 * math.pi * i) / n
        b["p"] += [h.center[0] + h.radius * math.cos(ang), h.center[1] + h.radius * math.sin(ang)]
        b["n"] += [-math.cos(ang), -math.sin(ang)]
        b["k"].append(-1.0 / h.radius)
        b["w"].append(2.0 * math.pi * h.radius / n)


def _star_knots(cx, cy, scale, n=20):
    xs, ys = [], []
    for i in range(n):
        ang = 2 * math.pi * (i / (n + 0.))
        xs.append(cx + scale * math.cos(ang) * (math.sin(5 * ang) + 4))
        ys.append(cy + scale * math.sin(ang) * (math.sin(5 * ang) + 4))
    return xs, ys


# --------------------------------------------------------------------------
# boundary data over node ranges (boundary.cpp:43-120)
# --------------------------------------------------------------------------
ONE_D = {"SINGLE_ELECTRON", "ALL_ONES", "ALL_NEG_ONES", "ALL_ZEROS", "EX3A"}


def bc_size(b: Boundary, bc: str) -> int:
    """set_boundary_values_size (boundary.cpp:11-40)."""
    return b.num_points if bc in ONE_D else 2 * b.num_points


def apply_bc(b: Boundary, f: np.ndarray, s: int, e: int, bc: str):
    xy, nm = b.xy(), b.normals.reshape(-1, 2)
    for p in range(s, e):
        x, y = xy[p]
        nx, ny = nm[p]
        if bc == "SINGLE_ELECTRON":
            f[p] = math.log(math.sqrt(math.pow(x + 3, 2) + math.pow(y + 2, 2))) / (2 * math.pi)
        elif bc == "ALL_ONES":
            f[p] = 1.0
        elif bc == "ALL_NEG_ONES":
            f[p] = -1.0
        elif bc == "ALL_ZEROS":
            f[p] = 0.0
        elif bc == "TANGENT_VEC":
            f[2 * p], f[2 * p + 1] = -ny, nx
        elif bc == "REVERSE_TANGENT_VEC":
            f[2 * p], f[2 * p + 1] = ny, -nx
        elif bc == "NORMAL_VEC":
            f[2 * p], f[2 * p + 1] = nx, ny
        elif bc == "REVERSE_NORMAL_VEC":
            f[2 * p], f[2 * p + 1] = -nx, -ny
        elif bc == "HORIZONTAL_VEC":
            f[2 * p], f[2 * p + 1] = 1.0, 0.0
        elif bc == "LEFT_TO_RIGHT_FLOW":
            f[2 * p], f[2 * p + 1] = (1.0 if (x < -0.99 or x > 1.99) else 0.0), 0.0
        elif bc == "NO_SLIP":
            f[2 * p], f[2 * p + 1] = 0.0, 0.0
        else:
            raise ValueError(bc)


def _whole(b: Boundary, bc: str):
    f = np.zeros(bc_size(b, bc))
    apply_bc(b, f, 0, b.num_points, bc)
    b.boundary_values = f
    return b


# --------------------------------------------------------------------------
# the reference's spline geometries
# --------------------------------------------------------------------------
def starfish(N: int, bc: str = "SINGLE_ELECTRON") -> Boundary:
    """CubicSpline::initialize (cubic_spline.cpp:29-50): 20 knots of
    0.5 + 0.375 (cos a, sin a)(sin 5a + 4), N / 20 nodes per interval.  With
    DEFAULT the reference applies SINGLE_ELECTRON over N nodes, more than the
    20 (N / 20) it made; here it is applied to the nodes it has."""
    nps = N // NUM_SPLINE_POINTS
    xs, ys = _star_knots(0.5, 0.5, 0.375)
    b = _new()
    xc, yc = get_cubics(xs, ys)
    interpolate(b, False, nps, xc, yc)
    out = _finish(b, [], nps * NUM_SPLINE_POINTS, "starfish")
    return _whole(out, "SINGLE_ELECTRON" if bc == "DEFAULT" else bc)


def ex1(N: int, bc: str = "DEFAULT", params=None) -> Boundary:
    """Ex1Boundary::initialize (ex1boundary.cpp:60-171): a square with wavy
    walls (24 knots), two small star holes and two circular holes."""
    params = [math.pi / 2.] if params is None else list(params)
    ospl, sspl = 24, 20
    onps = (2 * N // 3) // ospl
    snps = (N // 12) // sspl
    ncirc = N // 12
    th = params[0]
    holes = [Hole((0.5 + 0.3 * math.cos(math.pi + th), 0.5 + 0.3 * math.sin(math.pi + th)), 0.05,
                  sspl * snps),
             Hole((0.4, 0.5), 0.05, sspl * snps),
             Hole((0.6, 0.5), 0.05, ncirc),
             Hole((0.5 + 0.3 * math.cos(th), 0.5 + 0.3 * math.sin(th)), 0.05, ncirc)]
    xs, ys = [], []
    for i in range(3):
        xs += [i / 3.0, (i / 3.0) + (1.0 / 6.0)]
        ys += [0.0, 0.05]
    for i in range(3):
        xs += [1.0, 0.95]
        ys += [i / 3.0, (i / 3.0) + (1.0 / 6.0)]
    for i in range(3):
        xs += [1.0 - (i / 3.0), 1.0 - (1.0 / 6.0) - (i / 3.0)]
        ys += [1.0, 0.95]
    for i in range(3):
        xs += [0.0, 0.05]
        ys += [1.0 - (i / 3.0), 1.0 - (1.0 / 6.0) - (i / 3.0)]
    b = _new()
    xc, yc = get_cubics(xs, ys)
    interpolate(b, False, onps, xc, yc)
    for h in holes[:2]:
        sx, sy = _star_knots(h.center[0], h.center[1], 0.015)
        xc, yc = get_cubics(sx, sy)
        interpolate(b, True, snps, xc, yc)
    for h in holes[2:]:
        _circle_hole(b, h)
    out = _finish(b, holes, ospl * onps, "ex1")
    if bc != "DEFAULT":
        return _whole(out, bc)
    f = np.zeros(2 * out.num_points)
    b1 = ospl * onps
    b2 = b1 + sspl * snps
    b3 = b2 + sspl * snps
    b4 = b3 + ncirc
    b5 = b4 + ncirc
    for s, e, c in [(0, b1, "TANGENT_VEC"), (b1, b2, "REVERSE_NORMAL_VEC"), (b2, b3, "NORMAL_VEC"),
                    (b3, b4, "REVERSE_NORMAL_VEC"), (b4, b5, "NORMAL_VEC")]:
        apply_bc(out, f, s, e, c)
    out.boundary_values = f
    return out


def ex2(N: int, bc: str = "DEFAULT", params=None) -> Boundary:
    """Ex2Boundary::initialize (ex2boundary.cpp:34-92): a channel [-1, 2] x
    [0.25, 0.75] (28 knots) with eight circular holes at heights params[i]."""
    params = [0.5] * 8 if params is None else list(params)
    ospl = 28
    onps = (3 * N // 4) // ospl
    ncirc = (N // 4) // 8
    holes = [Hole((-0.7 + (i / 3.), params[i]), 0.05, ncirc) for i in range(8)]
    xs, ys = [], []
    for i in range(12):
        xs.append(-1 + 3 * (i / 12.0))
        ys.append(0.25)
    for i in range(2):
        xs.append(2.0)
        ys.append(0.25 + 0.5 * i / 2.0)
    for i in range(12):
        xs.append(2 - 3 * (i / 12.0))
        ys.append(0.75)
    for i in range(2):
        xs.append(-1.0)
        ys.append(0.75 - 0.5 * i / 2.0)
    b = _new()
    xc, yc = get_cubics(xs, ys)
    interpolate(b, False, onps, xc, yc)
    for h in holes:
        _circle_hole(b, h)
    out = _finish(b, holes, ospl * onps, "ex2")
    return _whole(out, "LEFT_TO_RIGHT_FLOW" if bc == "DEFAULT" else bc)


def ex3(N: int, bc: str = "EX3B", params=None) -> Boundary:
    """Ex3Boundary::initialize (ex3boundary.cpp:50-130): the starfish with
    two star holes on the circle of radius 0.2 (sin 5a + 4) at angles
    params[0], params[1]; EX3A (Laplace) or EX3B (Stokes) data."""
    params = [0.0, math.pi] if params is None else list(params)
    ospl, sspl = 20, 20
    onps = (2 * N // 3) // ospl
    snps = (N // 6) // sspl
    holes = []
    for a in params[:2]:
        x = 0.2 * math.cos(a) * (math.sin(5 * a) + 4)
        y = 0.2 * math.sin(a) * (math.sin(5 * a) + 4)
        holes.append(Hole((0.5 + x, 0.5 + y), 0.3, sspl * snps))
    b = _new()
    xs, ys = _star_knots(0.5, 0.5, 0.375)
    xc, yc = get_cubics(xs, ys)
    interpolate(b, False, onps, xc, yc)
    for h in holes:
        sx, sy = _star_knots(h.center[0], h.center[1], 0.06)
        xc, yc = get_cubics(sx, sy)
        interpolate(b, True, snps, xc, yc)
    out = _finish(b, holes, ospl * onps, "ex3")
    if bc not in ("EX3A", "EX3B"):
        if bc == "DEFAULT":
            raise ValueError("Ex3Boundary has no DEFAULT data")   # the reference exits
        return _whole(out, bc)
    b1 = ospl * onps
    b2 = b1 + sspl * snps
    b3 = b2 + sspl * snps
    if bc == "EX3A":
        f = np.zeros(out.num_points)
        parts = [(0, b1, "ALL_ZEROS"), (b1, b2, "ALL_ONES"), (b2, b3, "ALL_NEG_ONES")]
    else:
        f = np.zeros(2 * out.num_points)
        parts = [(0, b1, "HORIZONTAL_VEC"), (b1, b2, "REVERSE_NORMAL_VEC"), (b2, b3, "NORMAL_VEC")]
    for s, e, c in parts:
        apply_bc(out, f, s, e, c)
    out.boundary_values = f
    return out


# --------------------------------------------------------------------------
# is_in_domain rules
# --------------------------------------------------------------------------
def _cmod(a: int, m: int) -> int:
    """C's % (truncated toward zero)."""
    r = abs(a) % m
    return -r if a < 0 else r


def in_disc_domain(targets, center, radius, holes=(), eps=EPS_DOMAIN) -> np.ndarray:
    """Circle / Annulus / Donut: outside when |t - c| + eps > R or when
    |t - h| - eps < r for a hole (circle.cpp:34-42, annulus.cpp:66-78,
    donut.cpp:67-79)."""
    t = np.asarray(targets, np.float64).reshape(-1, 2)
    dx, dy = t[:, 0] - center[0], t[:, 1] - center[1]
    inside = ~(np.sqrt(dx * dx + dy * dy) + eps > radius)
    for h in holes:
        hx, hy = t[:, 0] - h.center[0], t[:, 1] - h.center[1]
        inside &= ~(np.sqrt(hx * hx + hy * hy) - eps < h.radius)
    return inside


def in_winding_domain(b: Boundary, targets, eps=EPS_DOMAIN) -> np.ndarray:
    """CubicBoundary::is_in_domain (cubic_boundary.cpp:182-230), vectorised
    over targets."""
    t = np.asarray(targets, np.float64).reshape(-1, 2)
    xy = b.xy()
    secs = [(0, b.num_outer)]
    o = b.num_outer
    for h in b.holes:
        secs.append((o, o + h.num_nodes))
        o += h.num_nodes
    inside = np.ones(t.shape[0], bool)
    for si, (s, e) in enumerate(secs):
        P = xy[s:e]
        Q = np.roll(P, -1, axis=0)
        w = np.zeros(t.shape[0], np.int64)
        for (px, py), qy in zip(P, Q[:, 1]):
            d = np.sqrt((t[:, 0] - px) ** 2 + (t[:, 1] - py) ** 2)
            inside &= ~(d < eps)
            right = px > t[:, 0]
            up = right & (py < t[:, 1]) & (qy > t[:, 1])
            down = right & ~up & (py > t[:, 1]) & (qy < t[:, 1])
            w += up.astype(np.int64) - down.astype(np.int64)
        rem = np.array([_cmod(int(v), 2) for v in w])
        inside &= (rem != 0) if si == 0 else (rem != 1)
    return inside


def in_star_domain(targets, amps, freqs, phases, eps=EPS_DOMAIN) -> np.ndarray:
    """StarCurve (no reference counterpart): |t| + eps < 1 + sum a_k cos(f_k theta + phi_k)."""
    t = np.asarray(targets, np.float64).reshape(-1, 2)
    th = np.arctan2(t[:, 1], t[:, 0])
    r = np.ones_like(th)
    for a, f, p in zip(amps, freqs, phases):
        r = r + a * np.cos(f * th + p)
    return np.sqrt(t[:, 0] ** 2 + t[:, 1] ** 2) + eps < r
