"""Problem assembly and end-to-end drivers.

TEST INFRASTRUCTURE ONLY (see the package docstring).

Restates ``kern_interp/linear_solve.cpp``: hole border matrices ``U``/``Psi``
(``:18-110``), ``linear_solve`` (``:113-123``), ``schur_solve``
(``:126-140``) and ``boundary_integral_solve`` (``:143-180``), plus the
reference test's residual check ``check_solve_err``
(``kern_interp_test.cpp:26-95``, without its ``err1``-twice slip at ``:78``).

Pinned by (``tests/test_oracle_skel.py``): the bordered dense system
``[K U; Psi -I]`` solved with LAPACK, and the analytic interior solutions.
"""
from __future__ import annotations

import math

import numpy as np

from . import kernel as kmod
from .skel import SkelFactorization
from .tree import QuadTree


def initialize_U(pde, holes, pts):
    """``initialize_U_mat`` (linear_solve.cpp:18-64)."""
    if not holes or pde == kmod.LAPLACE_NEUMANN:
        return np.zeros((0, 0))
    xy = np.asarray(pts, np.float64).reshape(-1, 2)
    if pde == kmod.LAPLACE:
        U = np.empty((xy.shape[0], len(holes)))
        for h, hole in enumerate(holes):
            rx = xy[:, 0] - hole.center[0]
            ry = xy[:, 1] - hole.center[1]
            U[:, h] = np.log(np.sqrt(rx * rx + ry * ry))
        return U
    U = np.zeros((2 * xy.shape[0], 3 * len(holes)))
    sc = 1.0 / (4 * math.pi)
    for h, hole in enumerate(holes):
        rx = xy[:, 0] - hole.center[0]
        ry = xy[:, 1] - hole.center[1]
        nr = np.sqrt(rx * rx + ry * ry)
        inv2 = 1.0 / nr ** 2
        U[0::2, 3 * h] = sc * (np.log(1 / nr) + inv2 * rx * rx)
        U[1::2, 3 * h] = sc * (inv2 * ry * rx)
        U[0::2, 3 * h + 1] = sc * (inv2 * rx * ry)
        U[1::2, 3 * h + 1] = sc * (np.log(1 / nr) + inv2 * ry * ry)
        U[0::2, 3 * h + 2] = ry * (sc / nr ** 2)
        U[1::2, 3 * h + 2] = -rx * (sc / nr ** 2)
    return U


def initialize_Psi(pde, holes, boundary):
    """``initialize_Psi_mat`` (linear_solve.cpp:67-110): hole membership by
    ``|x - c| < radius + 1e-8``, first matching hole wins."""
    if not holes or pde == kmod.LAPLACE_NEUMANN:
        return np.zeros((0, 0))
    xy = boundary.xy()
    w = boundary.weights
    npt = xy.shape[0]
    owner = np.full(npt, -1)
    for h in range(len(holes) - 1, -1, -1):
        c = holes[h].center
        d = np.sqrt((xy[:, 0] - c[0]) ** 2 + (xy[:, 1] - c[1]) ** 2)
        owner[d < holes[h].radius + 1e-8] = h
    if pde == kmod.LAPLACE:
        Psi = np.zeros((len(holes), npt))
        for i in np.nonzero(owner >= 0)[0]:
            Psi[owner[i], i] = w[i]
        return Psi
    Psi = np.zeros((3 * len(holes), 2 * npt))
    for i in np.nonzero(owner >= 0)[0]:
        h = owner[i]
        Psi[3 * h, 2 * i] = w[i]
        Psi[3 * h + 1, 2 * i + 1] = w[i]
        Psi[3 * h + 2, 2 * i] = w[i] * xy[i, 1]
        Psi[3 * h + 2, 2 * i + 1] = -w[i] * xy[i, 0]
    return Psi


def factor(boundary, pde, id_tol, leaf_dofs=32, fact_threads=1):
    """Tree build + ``skeletonize`` with the hole border set (linear_solve.cpp:154-163)."""
    sd = 2 if pde == kmod.STOKES else 1
    tree = QuadTree(boundary.points, sd, leaf_dofs)
    kern = kmod.Kernel(sd, pde, boundary)
    U = initialize_U(pde, boundary.holes, boundary.points)
    tree.U = U if U.size else None
    tree.Psi = initialize_Psi(pde, boundary.holes, boundary) if U.size else None
    fac = SkelFactorization(id_tol, fact_threads)
    fac.skeletonize(kern, tree)
    return tree, kern, fac


def linear_solve(fac, tree, f):
    """``linear_solve`` (linear_solve.cpp:113-123). Returns (mu, alpha or None)."""
    if tree.U is None:
        return fac.solve(tree, f), None
    return fac.multiply_connected_solve(tree, f)


def boundary_integral_solve(boundary, pde, id_tol, domain_points, leaf_dofs=32):
    """``boundary_integral_solve`` (linear_solve.cpp:143-180) without the
    ``is_in_domain`` masking of ``:168-178`` (callers pick interior targets)."""
    tree, kern, fac = factor(boundary, pde, id_tol, leaf_dofs)
    mu, alpha = linear_solve(fac, tree, boundary.boundary_values)
    kd = kmod.Kernel(kern.sd, pde, boundary, domain_points)
    sol = kd.forward() @ mu
    if alpha is not None:
        sol = sol + initialize_U(pde, boundary.holes, domain_points) @ alpha
    return sol, mu, alpha, tree, fac


def dense_operator(boundary, pde):
    """The full bordered operator ``[K U; Psi -I]`` (kern_interp_test.cpp:59-79)."""
    sd = 2 if pde == kmod.STOKES else 1
    kern = kmod.Kernel(sd, pde, boundary)
    nd = sd * boundary.num_points
    K = kern(np.arange(nd), np.arange(nd))
    U = initialize_U(pde, boundary.holes, boundary.points)
    if U.size == 0:
        return K
    Psi = initialize_Psi(pde, boundary.holes, boundary)
    h = U.shape[1]
    return np.block([[K, U], [Psi, -np.eye(h)]])


def backward_error(boundary, pde, mu, alpha=None):
    """Relative residual of the full system (kern_interp_test.cpp:26-95)."""
    A = dense_operator(boundary, pde)
    f = boundary.boundary_values
    if alpha is None:
        res = A @ mu - f
        return np.linalg.norm(res) / np.linalg.norm(f)
    x = np.concatenate([mu, alpha])
    rhs = np.concatenate([f, np.zeros(alpha.size)])
    return np.linalg.norm(A @ x - rhs) / np.linalg.norm(f)
