"""Pins for the oracle ID, factorization and solves.

* ID: exact rank on matrices with a designed singular-value gap, and the ID
  error bound |A P - A_s [I T]| <= c tol |A|;
* factor/solve: the reference's backward-error criterion
  |K mu - f| / |f| <= 10 id_tol (kern_interp_test.cpp:26-95) with a dense
  LAPACK operator, the circle's Sherman-Morrison closed form, and analytic
  interior solutions (kern_interp_test.cpp:297-316, 392-410).
"""
import math

import numpy as np
import pytest

from oracle import geometry as G
from oracle import kernel as K
from oracle import linear_solve as LS
from oracle.skel import interpolative_decomposition


@pytest.mark.parametrize("k", [1, 5, 17])
def test_id_rank_gap(k):
    rng = np.random.default_rng(k)
    m, n = 300, 40
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.concatenate([np.logspace(0, -2, k), np.full(n - k, 1e-10)])
    A = (U * s) @ V.T
    kk, perm, T = interpolative_decomposition(A, 1e-6)
    assert kk == k
    Ap = A[:, perm]
    err = np.linalg.norm(Ap - Ap[:, :k] @ np.hstack([np.eye(k), T]))
    assert err <= 1e-6 * np.linalg.norm(A)


def test_id_full_rank_is_not_compressed():
    rng = np.random.default_rng(0)
    kk, perm, T = interpolative_decomposition(rng.standard_normal((50, 20)), 1e-6)
    assert kk == 0 and perm is None


def _solve(b, pde, tol, leaf):
    tree, kern, fac = LS.factor(b, pde, tol, leaf)
    mu, alpha = LS.linear_solve(fac, tree, b.boundary_values)
    return mu, alpha


@pytest.mark.parametrize("name,make,bc,pde,leaf", [
    ("circle_laplace", lambda: G.circle(1024), G.bc_single_electron, K.LAPLACE, 32),
    ("star_laplace", lambda: G.star(4096), lambda b: G.bc_point_source(b, (3, 3)), K.LAPLACE, 64),
    ("annulus_laplace", lambda: G.annulus(1024), G.bc_single_electron, K.LAPLACE, 32),
    ("circle_stokes", lambda: G.circle(1024), G.bc_tangent, K.STOKES, 32),
])
def test_backward_error(name, make, bc, pde, leaf):
    b = make()
    b.boundary_values = bc(b)
    mu, alpha = _solve(b, pde, 1e-6, leaf)
    assert LS.backward_error(b, pde, mu, alpha) <= 10 * 1e-6


def test_circle_sherman_morrison():
    # K = I/2 + c 11^T on a circle, so K^{-1} = 2 I - 4c/(1 + 2 N c) 11^T.  The
    # far blocks are rank one, but the proxy rows make each ID a generic
    # rank-~10 truncation, so agreement is ~1e-10 (measured 3e-10), not 1e-16
    n = 2048
    b = G.circle(n)
    b.boundary_values = G.bc_single_electron(b)
    mu, _ = _solve(b, K.LAPLACE, 1e-6, 32)
    c = b.weights[0] / (4 * math.pi * 0.25)
    f = b.boundary_values
    exact = 2 * f - (4 * c / (1 + 2 * n * c)) * f.sum()
    assert np.linalg.norm(mu - exact) / np.linalg.norm(exact) < 1e-8


def test_laplace_circle_analytic():
    # kern_interp_test.cpp:392-410: N = 2^12, id_tol 1e-13, 20x20 grid, err <= 1e-12
    n = 4096
    b = G.circle(n)
    b.boundary_values = G.bc_single_electron(b)
    lo, hi = float(b.points.min()) - 0.0, float(b.points.max())
    g = np.linspace(lo, hi, 20)
    X, Y = np.meshgrid(g, g, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel()], 1)
    inside = np.hypot(pts[:, 0] - 0.5, pts[:, 1] - 0.5) + 1e-2 <= 0.25
    tgt = pts[inside].reshape(-1)
    sol, mu, _, _, _ = LS.boundary_integral_solve(b, K.LAPLACE, 1e-13, tgt, 32)
    xy = tgt.reshape(-1, 2)
    ex = np.log(np.sqrt((xy[:, 0] + 3) ** 2 + (xy[:, 1] + 2) ** 2)) / (2 * np.pi)
    assert np.linalg.norm(sol - ex) / np.linalg.norm(ex) <= 10 * 1e-13


def test_laplace_annulus_analytic():
    # kern_interp_test.cpp:413-430: the bordered (hole) system
    b = G.annulus(2048)
    b.boundary_values = G.bc_single_electron(b)
    tgt = []
    for x in np.linspace(-0.5, 1.5, 20):
        for y in np.linspace(-0.5, 1.5, 20):
            if np.hypot(x - 0.5, y - 0.5) + 1e-2 > 1:
                continue
            if any(np.hypot(x - h.center[0], y - h.center[1]) - 1e-2 < h.radius for h in b.holes):
                continue
            tgt.append((x, y))
    tgt = np.array(tgt).reshape(-1)
    sol, mu, alpha, _, _ = LS.boundary_integral_solve(b, K.LAPLACE, 1e-13, tgt, 32)
    xy = tgt.reshape(-1, 2)
    ex = np.log(np.sqrt((xy[:, 0] + 3) ** 2 + (xy[:, 1] + 2) ** 2)) / (2 * np.pi)
    assert np.linalg.norm(sol - ex) / np.linalg.norm(ex) <= 1e-10


def test_star_interior_error_config():
    # the factor/solve workload at its CPU size: star N=4096, leaf 64, tol 1e-6,
    # point source at (3, 3), 1000 seeded targets with r < 0.6
    b = G.star(4096)
    b.boundary_values = G.bc_point_source(b, (3.0, 3.0))
    tgt = G.interior_disc_targets(1000, 1234)
    sol, mu, _, tree, fac = LS.boundary_integral_solve(b, K.LAPLACE, 1e-6, tgt, 64)
    xy = tgt.reshape(-1, 2)
    ex = np.log(np.sqrt((xy[:, 0] - 3) ** 2 + (xy[:, 1] - 3) ** 2)) / (2 * np.pi)
    assert np.linalg.norm(sol - ex) / np.linalg.norm(ex) <= 1e-6
    dense = LS.dense_operator(b, K.LAPLACE)
    mu_d = np.linalg.solve(dense, b.boundary_values)
    assert np.linalg.norm(mu - mu_d) / np.linalg.norm(mu_d) <= 1e-5
