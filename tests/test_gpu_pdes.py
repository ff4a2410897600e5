"""Stokes (2x2 DOF blocks, kernel.cpp:47-96, 268-307) and Laplace Neumann
(kernel.cpp:24-33) through the GPU factor/solve, against the oracle and the
reference's backward-error criterion (kern_interp_test.cpp:26-95)."""
import numpy as np
import pytest

import kigpu
from oracle import geometry as G
from oracle import kernel as K
from oracle import linear_solve as LS

from helpers import gpu_tree


def factor_both(ctx, b, pde, leaf, tol=1e-6):
    sd = 2 if pde == K.STOKES else 1
    otree, okern, ofac = LS.factor(b, pde, tol, leaf)
    t = gpu_tree(ctx, b, sd, leaf)
    t.factor(pde, tol, kigpu.FP64)
    rank, active = t.ranks()
    orank = np.array([len(n.skel) if n.compressed else -1 for n in otree.all_nodes()])
    return otree, ofac, t, rank, orank


@pytest.mark.gpu
def test_stokes_star_parity(ctx):
    print("[pde] stokes star N=2048 leaf 64", flush=True)
    b = G.star(2048)
    b.boundary_values = G.bc_tangent(b)
    otree, ofac, t, rank, orank = factor_both(ctx, b, K.STOKES, 64)
    assert np.array_equal(rank, orank)
    so, s, ro, r = t.lists()
    for i, nd in enumerate(otree.all_nodes()):
        if nd.compressed:
            assert list(s[so[i]:so[i + 1]]) == nd.skel
    mu = t.solve(b.boundary_values)
    mu_ref = ofac.solve(otree, b.boundary_values)
    assert np.linalg.norm(mu - mu_ref) / np.linalg.norm(mu_ref) < 1e-9
    assert LS.backward_error(b, K.STOKES, mu) <= 10 * 1e-6


@pytest.mark.gpu
def test_stokes_circle_backward_error(ctx):
    # kern_interp_test.cpp StokesCircleBackwardError (N = 2^10, tangent data)
    print("[pde] stokes circle N=1024 leaf 32", flush=True)
    b = G.circle(1024)
    b.boundary_values = G.bc_tangent(b)
    otree, ofac, t, rank, orank = factor_both(ctx, b, K.STOKES, 32)
    assert np.array_equal(rank, orank)
    mu = t.solve(b.boundary_values)
    assert LS.backward_error(b, K.STOKES, mu) <= 10 * 1e-6
    # rigid rotation inside (kern_interp_test.cpp:359-389)
    tgt = G.interior_disc_targets(50, 4, radius=0.15) + 0.5
    u = t.forward(tgt, mu).reshape(-1, 2)
    rr = tgt.reshape(-1, 2) - 0.5
    truth = 4.0 * np.stack([-rr[:, 1], rr[:, 0]], 1)
    assert np.linalg.norm(u - truth) / np.linalg.norm(truth) < 1e-6


@pytest.mark.gpu
def test_neumann_circle_backward_error(ctx):
    # kern_interp_test.cpp LaplaceNeumannCircleBackwardError (ALL_ONES data)
    print("[pde] laplace neumann circle N=1024 leaf 32", flush=True)
    b = G.circle(1024)
    b.boundary_values = np.ones(1024)
    otree, ofac, t, rank, orank = factor_both(ctx, b, K.LAPLACE_NEUMANN, 32)
    assert np.array_equal(rank, orank)
    mu = t.solve(b.boundary_values)
    assert LS.backward_error(b, K.LAPLACE_NEUMANN, mu) <= 10 * 1e-6


@pytest.mark.gpu
def test_neumann_star_parity(ctx):
    print("[pde] laplace neumann star N=4096 leaf 64", flush=True)
    b = G.star(4096)
    b.boundary_values = G.bc_point_source(b, (3.0, 3.0))
    otree, ofac, t, rank, orank = factor_both(ctx, b, K.LAPLACE_NEUMANN, 64)
    assert np.array_equal(rank, orank)
    mu = t.solve(b.boundary_values)
    mu_ref = ofac.solve(otree, b.boundary_values)
    assert np.linalg.norm(mu - mu_ref) / np.linalg.norm(mu_ref) < 1e-9
