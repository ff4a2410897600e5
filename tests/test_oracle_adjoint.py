"""Oracle pins for the adjoint (CPU): complex-step gradient vs finite
differences, the transposed solve vs the adjoint identity and the dense
K^{-T}, and the host-side geometry Jacobian of the shape module."""
import numpy as np

from oracle import adjoint as A
from oracle import geometry as G
from oracle import kernel as K
from oracle import linear_solve as LS

from kigpu.shape import FourierStar


def coeffs(m=16, seed=3, scale=0.05):
    return scale * (2 * G.uniform01(seed, m) - 1)


def test_param_star_zero_is_star():
    g = A.param_star(1024, np.zeros(16))
    b = G.star(1024)
    xy = b.points.reshape(-1, 2)
    assert np.array_equal(g["x"], xy[:, 0]) and np.array_equal(g["y"], xy[:, 1])
    assert np.allclose(g["w"], b.weights, rtol=1e-15, atol=0)
    assert np.allclose(g["kappa"], b.curvatures, rtol=1e-13, atol=1e-13)


def test_complex_step_matches_central_differences():
    # the complex step and a central difference (h = 1e-6) agree to O(h^2)
    c = coeffs()
    tgt = G.interior_disc_targets(64, 11, radius=0.5)
    cs = A.complex_step_grad(256, c, tgt)
    fd = A.fd_grad(256, c, tgt, h=1e-6)
    assert np.linalg.norm(cs - fd) <= 1e-6 * np.linalg.norm(cs)


def test_shape_jacobian_matches_complex_step():
    # kigpu.shape.FourierStar.jacobian (analytic chain rule) vs the complex
    # step through the oracle's discretisation
    n, c = 512, coeffs()
    star = FourierStar(n, c.size)
    J = star.jacobian(c)
    geo = star.geometry(c)
    g0 = A.param_star(n, c)
    for key, got in zip(("x", "y", "nx", "ny", "w", "kappa"), geo):
        assert np.allclose(got, g0[key], rtol=1e-14, atol=1e-14)
    for k in (0, 5, 15):
        cc = c.astype(np.complex128)
        cc[k] += 1e-30j
        gk = A.param_star(n, cc)
        for q, key in enumerate(("x", "y", "nx", "ny", "w", "kappa")):
            ref = np.imag(gk[key]) / 1e-30
            assert np.allclose(J[:, q, k], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_solve_transpose_adjoint_identity():
    b = G.star(2048)
    b.boundary_values = G.bc_point_source(b)
    tree, kern, fac = LS.factor(b, K.LAPLACE, 1e-6, 64)
    rng = np.random.default_rng(1)
    g, v = rng.standard_normal(2048), rng.standard_normal(2048)
    lhs = fac.solve_transpose(tree, g) @ v
    rhs = g @ fac.solve(tree, v)
    assert abs(lhs - rhs) <= 1e-12 * abs(rhs) * 10


def test_solve_transpose_tight_matches_dense():
    b = G.star(2048)
    b.boundary_values = G.bc_point_source(b)
    tree, kern, fac = LS.factor(b, K.LAPLACE, 1e-12, 64)
    Kd = kern(np.arange(2048), np.arange(2048))
    g = np.random.default_rng(2).standard_normal(2048)
    ref = np.linalg.solve(Kd.T, g)
    got = fac.solve_transpose(tree, g)
    assert np.linalg.norm(got - ref) <= 1e-9 * np.linalg.norm(ref)
