"""Pins for the oracle's kernel entries, proxy rows and forward matrix.

None of these re-types the entry formulas: each checks a property the
mathematics fixes (closed form on a circle, Gauss' law, the far-field ID
through the proxy ring, a rigid rotation for Stokes)."""
import math

import numpy as np
import pytest

from oracle import geometry as G
from oracle import kernel as K
from oracle import linear_solve as LS
from oracle.skel import interpolative_decomposition
from oracle.tree import QuadTree


def test_circle_closed_form():
    # on a circle of radius R, (r . n_s)/|r|^2 = -1/(2R) for every pair, so the
    # Laplace double layer is w/(4 pi R) off the diagonal and 1/2 + w/(4 pi R) on it
    b = G.circle(256)
    R = 0.25
    kern = K.Kernel(1, K.LAPLACE, b)
    A = kern(np.arange(256), np.arange(256))
    c = b.weights[0] / (4 * math.pi * R)
    off = A[~np.eye(256, dtype=bool)]
    # node coordinates carry ~1e-16 rounding from cos/sin; for neighbors
    # r . n_s ~ |r|^2 / (2R) amplifies it to ~1e-12 relative
    assert np.max(np.abs(off / c - 1)) < 1e-11
    assert np.allclose(np.diag(A), 0.5 + c, rtol=1e-14, atol=0)


@pytest.mark.parametrize("make", [lambda: G.star(2048), lambda: G.random_star(2048, 3)])
def test_gauss_law(make):
    # double layer of a constant density: 1 at interior targets, and with the
    # 1/2 jump and curvature-corrected diagonal also 1 on the boundary
    b = make()
    kern = K.Kernel(1, K.LAPLACE, b)
    idx = np.arange(0, b.num_points, 7)
    rows = kern(idx, np.arange(b.num_points)).sum(axis=1)
    assert np.max(np.abs(rows - 1.0)) < 1e-9
    tgt = G.interior_disc_targets(50, 8, radius=0.5)
    kd = K.Kernel(1, K.LAPLACE, b, tgt)
    assert np.max(np.abs(kd.forward().sum(axis=1) - 1.0)) < 1e-9


def test_proxy_ring_represents_far_field():
    # The ID of [near rows; proxy rows] of a box must interpolate the box's
    # interaction with every far source/target (Green's representation via the
    # proxy circle).  A wrong proxy entry breaks this at the 1e-6 level.
    b = G.star(4096)
    tree = QuadTree(b.points, 1, 64)
    kern = K.Kernel(1, K.LAPLACE, b)
    tree.refresh_level(len(tree.levels) - 1)
    leaf = max(tree.levels[-1], key=lambda n: len(n.active_box))
    A = kern.id_matrix(tree, leaf)
    k, perm, T = interpolative_decomposition(A, 1e-10)
    assert k > 0
    box = np.asarray(leaf.active_box)
    cx = leaf.corners[0] + leaf.side / 2
    cy = leaf.corners[1] + leaf.side / 2
    xy = b.xy()
    far = np.nonzero(np.hypot(xy[:, 0] - cx, xy[:, 1] - cy) > 1.6 * leaf.side)[0]
    # rows: far targets <- box, and box -> far sources (transposed)
    F = np.vstack([kern(far, box), kern(box, far).T])
    Fp = F[:, perm]
    approx = Fp[:, :k] @ np.hstack([np.eye(k), T])
    assert np.linalg.norm(approx - Fp) <= 1e-7 * np.linalg.norm(F)


def test_stokes_rigid_rotation():
    # kern_interp_test.cpp:359-389: tangent data on the circle gives the rigid
    # rotation u = 4 (-(y - 0.5), x - 0.5) inside
    b = G.circle(512)
    b.boundary_values = G.bc_tangent(b)
    A = LS.dense_operator(b, K.STOKES)
    mu = np.linalg.solve(A, b.boundary_values)
    tgt = (G.interior_disc_targets(40, 2, radius=0.15) + 0.5)
    kd = K.Kernel(2, K.STOKES, b, tgt)
    u = (kd.forward() @ mu).reshape(-1, 2)
    r = tgt.reshape(-1, 2) - 0.5
    truth = 4.0 * np.stack([-r[:, 1], r[:, 0]], 1)
    assert np.linalg.norm(u - truth) / np.linalg.norm(truth) < 1e-10


def test_neumann_circle_closed_form():
    # LAPLACE_NEUMANN on a circle of radius R (kernel.cpp:23-33): with
    # r = t - s and n_t = (t - c)/R, (r . n_t)/|r|^2 = 1/(2R) for every pair,
    # so the off-diagonal entries are w + w/(4 pi R) and the diagonal
    # (kappa = 1/R) w - 1/2 + w/(4 pi R): K = -I/2 + (w + w/(4 pi R)) 1 1^T
    b = G.circle(256)
    R = 0.25
    w = b.weights[0]
    kern = K.Kernel(1, K.LAPLACE_NEUMANN, b)
    A = kern(np.arange(256), np.arange(256))
    c = w + w / (4 * math.pi * R)
    ref = -0.5 * np.eye(256) + c
    assert np.max(np.abs(A - ref)) <= 1e-11 * c


def test_neumann_solve_sherman_morrison():
    # the factor/solve of the Neumann kernel on the circle against the inverse
    # of -I/2 + c 1 1^T: (1/a)(I - c 1 1^T / (a + c N)), a = -1/2
    n = 1024
    b = G.circle(n)
    b.boundary_values = G.bc_single_electron(b)
    R = 0.25
    w = b.weights[0]
    c = w + w / (4 * math.pi * R)
    tree, kern, fac = LS.factor(b, K.LAPLACE_NEUMANN, 1e-12, 32)
    mu, _ = LS.linear_solve(fac, tree, b.boundary_values)
    f = b.boundary_values
    a = -0.5
    exact = (f - c * f.sum() / (a + c * n)) / a
    assert np.linalg.norm(mu - exact) / np.linalg.norm(exact) < 1e-9


def test_neumann_forward_single_layer_of_constant():
    # the Neumann forward entries w/(2 pi) log|x - s| (kernel.cpp:24-27) sum a
    # trapezoidal single layer; for density 1 on a circle of radius R the
    # potential inside is R log R (mean value of log|x - s| over the circle)
    b = G.circle(512)
    R = 0.25
    tgt = G.interior_disc_targets(30, 5, radius=0.15) + 0.5
    kd = K.Kernel(1, K.LAPLACE_NEUMANN, b, tgt)
    u = kd.forward() @ np.ones(512)
    assert np.max(np.abs(u - R * math.log(R))) < 1e-12
