"""Adjoint path on the GPU: transposed solve, transposed applies, shape
sensitivities and the full shape gradient, against the oracle.

Bars: transposed solve within 1e-9 of the oracle's F^{-T} g (same compressed
operator, rounding only) and the adjoint identity to 1e-12; transposed
applies to 1e-12 of the dense products; sensitivities to 1e-10 of the
complex-step directional derivative of the dense bilinear forms; the
gradient within 1e-7 of the dense complex-step gradient (FP32 factor refined
to 1e-13 against the FP64 kernel, so both differentiate the same
uncompressed discretisation); at N = 32768 the gradient along a random
direction agrees with a central difference (h = 1e-6) of the GPU loss to
1e-5."""
import numpy as np
import pytest

import kigpu
from kigpu import shape
from oracle import adjoint as A
from oracle import geometry as G
from oracle import kernel as K
from oracle import linear_solve as LS

from helpers import gpu_tree, star_problem


def coeffs(m=16, seed=3, scale=0.02):
    return scale * (2 * G.uniform01(seed, m) - 1)


@pytest.mark.gpu
def test_solve_transpose_parity(ctx):
    print("[adjoint] transposed solve star N=4096 FP64", flush=True)
    b = star_problem(4096)
    otree, okern, ofac = LS.factor(b, K.LAPLACE, 1e-6, 64)
    t = gpu_tree(ctx, b)
    t.factor(K.LAPLACE, 1e-6, kigpu.FP64)
    rng = np.random.default_rng(4)
    g, v = rng.standard_normal((2, 4096))
    lam = t.solve_transpose(g)
    ref = ofac.solve_transpose(otree, g)
    assert np.linalg.norm(lam - ref) <= 1e-9 * np.linalg.norm(ref)
    x = t.solve(v)
    assert abs(lam @ v - g @ x) <= 1e-12 * abs(g @ x) * 10
    G2 = rng.standard_normal((4096, 3))
    L2 = t.solve_transpose(G2)
    for j in range(3):
        assert np.allclose(L2[:, j], t.solve_transpose(G2[:, j]), rtol=0,
                           atol=1e-13 * np.abs(L2[:, j]).max())


@pytest.mark.gpu
def test_solve_transpose_fp32_refined(ctx):
    print("[adjoint] transposed solve star N=4096 FP32 + refinement", flush=True)
    b = star_problem(4096)
    t = gpu_tree(ctx, b)
    t.factor(K.LAPLACE, 1e-6, kigpu.FP32_REFINE)
    g = np.random.default_rng(5).standard_normal(4096)
    lam = t.solve_transpose(g, refine_iters=10, rel_residual=1e-12)
    assert t.last_info["rel_residual"] <= 1e-12
    Kd = K.Kernel(1, K.LAPLACE, b)(np.arange(4096), np.arange(4096))
    assert np.linalg.norm(Kd.T @ lam - g) <= 2e-12 * np.linalg.norm(g)


@pytest.mark.gpu
@pytest.mark.parametrize("pde", [K.LAPLACE, K.LAPLACE_NEUMANN, K.STOKES])
def test_transposed_applies(ctx, pde):
    print(f"[adjoint] transposed applies pde {pde}", flush=True)
    b = G.circle(512) if pde == K.STOKES else G.star(1024)
    sd = 2 if pde == K.STOKES else 1
    t = gpu_tree(ctx, b, sd, 32)
    nd = sd * b.num_points
    rng = np.random.default_rng(6)
    lam = rng.standard_normal((nd, 2))
    Kd = K.Kernel(sd, pde, b)(np.arange(nd), np.arange(nd))
    y = t.apply_transpose(lam, pde)
    ref = Kd.T @ lam
    assert np.allclose(y, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    tgt = G.interior_disc_targets(300, 3) * 0.2 + (0.5 if pde == K.STOKES else 0.0)
    F = K.Kernel(sd, pde, b, tgt).forward()
    v = rng.standard_normal((F.shape[0], 2))
    y = t.forward_transpose(tgt, v, pde)
    ref = F.T @ v
    assert np.allclose(y, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def _soa(g):
    return {k: np.asarray(g[k]) for k in ("x", "y", "nx", "ny", "w", "kappa")}


@pytest.mark.gpu
def test_sensitivities_match_complex_step(ctx):
    print("[adjoint] sensitivities N=1024", flush=True)
    n = 1024
    g = _soa(A.param_star(n, coeffs()))
    pts = np.stack([g["x"], g["y"]], 1).reshape(-1)
    nrm = np.stack([g["nx"], g["ny"]], 1).reshape(-1)
    t = ctx.build_tree(pts, nrm, g["w"], g["kappa"], 1, 32)
    rng = np.random.default_rng(7)
    lam, mu = rng.standard_normal((2, n))
    dg = {k: rng.standard_normal(n) for k in g}
    order = ("x", "y", "nx", "ny", "w", "kappa")
    D = np.stack([dg[k] for k in order], 1)
    sK = t.kernel_sensitivity(lam, mu)
    ref = A.geometry_directional(g, dg, lambda gc: lam @ (A.dense_K(gc) @ mu))
    assert abs(np.sum(sK * D) - ref) <= 1e-10 * np.abs(sK * D).sum()
    tgt = G.interior_disc_targets(200, 9, radius=0.3)
    v = rng.standard_normal(200)
    sF = t.forward_sensitivity(tgt, v, mu)
    ref = A.geometry_directional(g, dg, lambda gc: v @ (A.dense_F(gc, tgt) @ mu))
    assert abs(np.sum(sF * D) - ref) <= 1e-10 * np.abs(sF * D).sum()
    t.close()


@pytest.mark.gpu
def test_shape_gradient_matches_complex_step(ctx):
    print("[adjoint] shape gradient N=2048, 16 coefficients", flush=True)
    n, c = 2048, coeffs()
    tgt = G.interior_disc_targets(128, 21, radius=0.3)
    star = shape.FourierStar(n, c.size)
    L, grad, info = shape.loss_and_gradient(ctx, star, c, tgt, id_tol=1e-10,
                                            precision=kigpu.FP32_REFINE, refine_iters=10,
                                            rel_residual=1e-13)
    ref = A.complex_step_grad(n, c, tgt)
    print(f"  |grad| {np.linalg.norm(ref):.3e} rel diff "
          f"{np.linalg.norm(grad - ref) / np.linalg.norm(ref):.3e}", flush=True)
    assert abs(L - A.dense_loss(n, c, tgt)) <= 1e-11 * L
    assert np.linalg.norm(grad - ref) <= 1e-7 * np.linalg.norm(ref)


@pytest.mark.gpu
def test_shape_gradient_n32768_directional(ctx):
    # N = 32768, 16 coefficients, 1024 targets: grad . d against a central
    # difference of the GPU loss itself (h = 1e-6)
    print("[adjoint] shape gradient N=32768 directional check", flush=True)
    n, c = 32768, coeffs()
    tgt = G.interior_disc_targets(1024, 33, radius=0.3)
    star = shape.FourierStar(n, c.size)
    kw = dict(id_tol=1e-10, precision=kigpu.FP32_REFINE, refine_iters=10, rel_residual=1e-13)
    L, grad, info = shape.loss_and_gradient(ctx, star, c, tgt, **kw)
    d = 2 * G.uniform01(44, c.size) - 1
    h = 1e-6
    Lp, _, _ = shape.loss_and_gradient(ctx, star, c + h * d, tgt, **kw)
    Lm, _, _ = shape.loss_and_gradient(ctx, star, c - h * d, tgt, **kw)
    fd = (Lp - Lm) / (2 * h)
    print(f"  grad.d {grad @ d:.10e} fd {fd:.10e}", flush=True)
    assert abs(grad @ d - fd) <= 1e-5 * abs(fd)


@pytest.mark.gpu
def test_shape_loss_autograd(ctx):
    # the torch.autograd wrapper: backward leaves the adjoint gradient of the
    # same call in c.grad (scaled by the incoming gradient), against the
    # dense complex-step gradient
    import torch
    print("[adjoint] torch.autograd wrapper, shape gradient N=2048", flush=True)
    n, c = 2048, coeffs()
    tgt = G.interior_disc_targets(128, 21, radius=0.3)
    star = shape.FourierStar(n, c.size)
    ct = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    L = shape.shape_loss(ctx, star, ct, tgt, id_tol=1e-10, precision=kigpu.FP32_REFINE,
                         refine_iters=10, rel_residual=1e-13)
    (2.0 * L).backward()
    ref = A.complex_step_grad(n, c, tgt)
    assert abs(L.item() - A.dense_loss(n, c, tgt)) <= 1e-11 * L.item()
    assert np.linalg.norm(ct.grad.numpy() / 2.0 - ref) <= 1e-7 * np.linalg.norm(ref)
