"""Adjoint shape gradients for a Fourier-parameterised star (Laplace double layer).

No counterpart in the reference, whose experiments differentiate by finite
differences of whole solves (``experiments/experiment_three.cpp:22-56,
160-171``).  The curve is

    r(t) = 1 + 0.3 cos 5t + sum_k (c_{2k} cos (k+1)t + c_{2k+1} sin (k+1)t),

discretised with N trapezoidal nodes (outward normal, weight |tau| 2pi/N,
curvature (x'y'' - x''y')/|tau|^3).  For the loss L = sum_j u(x_j)^2 at fixed
domain targets, with u = F mu, K mu = f and boundary data f(x, y) (default
x^2 + x y: data that is not the trace of a harmonic function, so that u and
L depend on the shape):

    dL/dc = [s_F - s_K] . dG/dc + lam^T df/dc,     lam = K^{-T} F^T (2u),

where G = (x, y, nx, ny, w, kappa) per node, s_K = d(lam^T K mu)/dG and
s_F = d(v^T F mu)/dG (v = 2u) come from the GPU, and dG/dc is the exact
derivative of the discretisation below.  Everything O(N^2) or heavier
(factor, solves, applies, sensitivities) runs on the device; the host does
the O(N m) chain rule.
"""
from __future__ import annotations

import numpy as np

from . import FP32_REFINE, FP64, LAPLACE

TWO_PI = 2.0 * np.pi


class FourierStar:
    """Nodes of the parameterised star and their derivatives in the coefficients."""

    def __init__(self, n: int, ncoef: int = 16):
        if ncoef % 2:
            raise ValueError("ncoef must be even (cos/sin pairs)")
        self.n = n
        self.ncoef = ncoef
        self.t = TWO_PI * np.arange(n) / n

    def _radius(self, c):
        t = self.t
        R = 1.0 + 0.3 * np.cos(5 * t)
        P = -1.5 * np.sin(5 * t)
        Q = -7.5 * np.cos(5 * t)
        # basis b_k(t) and its first two derivatives, (ncoef, n)
        B = np.zeros((self.ncoef, self.n))
        dB = np.zeros_like(B)
        ddB = np.zeros_like(B)
        for k in range(self.ncoef // 2):
            f = k + 1.0
            cs, sn = np.cos(f * t), np.sin(f * t)
            B[2 * k], B[2 * k + 1] = cs, sn
            dB[2 * k], dB[2 * k + 1] = -f * sn, f * cs
            ddB[2 * k], ddB[2 * k + 1] = -f * f * cs, -f * f * sn
        c = np.asarray(c, dtype=np.float64)
        return R + c @ B, P + c @ dB, Q + c @ ddB, B, dB, ddB

    def geometry(self, c):
        """(x, y, nx, ny, w, kappa), each of length n."""
        R, P, Q, _, _, _ = self._radius(c)
        ct, st = np.cos(self.t), np.sin(self.t)
        tx, ty = P * ct - R * st, P * st + R * ct
        s = np.sqrt(tx * tx + ty * ty)
        num = R * R + 2 * P * P - R * Q
        return (R * ct, R * st, ty / s, -tx / s, s * TWO_PI / self.n, num / s ** 3)

    def jacobian(self, c):
        """dG/dc as an (n, 6, ncoef) array, G = (x, y, nx, ny, w, kappa)."""
        R, P, Q, B, dB, ddB = self._radius(c)
        ct, st = np.cos(self.t), np.sin(self.t)
        tx, ty = P * ct - R * st, P * st + R * ct
        s2 = tx * tx + ty * ty          # = R^2 + P^2
        s = np.sqrt(s2)
        s3 = s2 * s
        num = R * R + 2 * P * P - R * Q
        # partials of each node quantity in (R, P, Q)
        d = np.zeros((self.n, 6, 3))
        d[:, 0, 0] = ct
        d[:, 1, 0] = st
        d[:, 2, 0] = ct / s - ty * R / s3
        d[:, 2, 1] = st / s - ty * P / s3
        d[:, 3, 0] = st / s + tx * R / s3
        d[:, 3, 1] = -ct / s + tx * P / s3
        d[:, 4, 0] = TWO_PI / self.n * R / s
        d[:, 4, 1] = TWO_PI / self.n * P / s
        d[:, 5, 0] = (2 * R - Q) / s3 - 3 * num * R / (s3 * s2)
        d[:, 5, 1] = 4 * P / s3 - 3 * num * P / (s3 * s2)
        d[:, 5, 2] = -R / s3
        basis = np.stack([B.T, dB.T, ddB.T], 1)   # (n, 3, ncoef)
        return np.einsum("nqa,nac->nqc", d, basis)


def quadratic_bc(x, y):
    """f = x^2 + x y and its gradient (df/dx, df/dy)."""
    return x * x + x * y, 2 * x + y, x


def loss_and_gradient(ctx, star: FourierStar, c, targets, bc=quadratic_bc, id_tol=1e-10,
                      leaf_dofs=64, precision=FP64, refine_iters=0, rel_residual=1e-12):
    """L = sum_j u(x_j)^2 and dL/dc through one factorization on the GPU.

    Returns (L, grad, info) where info holds the solve residuals.  With
    precision FP32_REFINE and refine_iters > 0 both solves are refined against
    the FP64 kernel, so the gradient is that of the true (uncompressed)
    discretisation; with FP64 it is that of the compressed operator, accurate
    to about id_tol.
    """
    x, y, nx, ny, w, kap = star.geometry(c)
    pts = np.stack([x, y], 1).reshape(-1)
    nrm = np.stack([nx, ny], 1).reshape(-1)
    tree = ctx.build_tree(pts, nrm, w, kap, 1, leaf_dofs)
    try:
        tree.factor(LAPLACE, id_tol, precision)
        f, fx, fy = bc(x, y)
        it = refine_iters if precision == FP32_REFINE else 0
        mu = tree.solve(f, refine_iters=it, rel_residual=rel_residual)
        info = {"solve": dict(tree.last_info)}
        u = tree.forward(targets, mu)
        v = 2.0 * u
        lam = tree.solve_transpose(tree.forward_transpose(targets, v), refine_iters=it,
                                   rel_residual=rel_residual)
        info["solve_transpose"] = dict(tree.last_info)
        sK = tree.kernel_sensitivity(lam, mu)
        sF = tree.forward_sensitivity(targets, v, mu)
    finally:
        tree.close()
    J = star.jacobian(c)
    grad = np.einsum("nq,nqc->c", sF - sK, J)
    # lam^T df/dc through the node positions
    grad += np.einsum("n,nc->c", lam * fx, J[:, 0, :])
    grad += np.einsum("n,nc->c", lam * fy, J[:, 1, :])
    return float(np.sum(u * u)), grad, info


def shape_loss(ctx, star: FourierStar, c, targets, **kw):
    """L(c) as a torch scalar whose backward pass is the adjoint gradient.

    ``c`` is a 1-D float64 tensor of the star's coefficients (any device; the
    computation itself runs through the library on its own context and
    stream).  The forward pass calls :func:`loss_and_gradient` once and keeps
    dL/dc; ``L.backward()`` then scales it by the incoming gradient, so no
    second factorization runs.  Keyword arguments go to
    :func:`loss_and_gradient`.  torch is imported here, not by the package.
    """
    import torch

    class _ShapeLoss(torch.autograd.Function):
        @staticmethod
        def forward(fctx, coeffs):
            L, g, info = loss_and_gradient(ctx, star, coeffs.detach().cpu().double().numpy(),
                                           targets, **kw)
            fctx.save_for_backward(torch.as_tensor(g, dtype=coeffs.dtype, device=coeffs.device))
            fctx.info = info
            return torch.tensor(L, dtype=coeffs.dtype, device=coeffs.device)

        @staticmethod
        def backward(fctx, gout):
            (g,) = fctx.saved_tensors
            return gout * g

    return _ShapeLoss.apply(c)
