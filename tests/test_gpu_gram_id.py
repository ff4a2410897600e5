"""The ID from the Gram matrix (kigpu_set_id_method; the FP32 factorization's
default) against the oracle's dgeqp3 ID (ki_mat.cpp:505-541) and against the
GPU's pivoted-QR ID.

In exact arithmetic the pivots, the rank and T = R11^{-1} R12 of a pivoted
QR depend on A only through A^T A, so the pivoted Cholesky of A^T A gives
the QR's ID.  In FP64 the remaining column norms it compares carry an
absolute error ~ eps * |A_j|^2, so boxes whose pivots or rank FP64 cannot
separate (near ties) may come out differently: the bar is the same rank and
ordered skeleton on all but 1% of the boxes, the density within 1e-8 of the
oracle's and a backward error <= 10 id_tol (kern_interp_test.cpp:90-93).
The FP32 factorization refines against the FP64 kernel apply, so with
either ID it must reach the same refined density.
"""
import numpy as np
import pytest

import kigpu
from oracle import geometry as G
from oracle import kernel as K
from oracle import linear_solve as LS

from helpers import gpu_tree, pde_sd, star_problem


def compare_with_oracle(ctx, b, pde, leaf, tol):
    sd = pde_sd(pde)
    otree, okern, ofac = LS.factor(b, pde, tol, leaf)
    ctx.set_id_method(2)
    try:
        t = gpu_tree(ctx, b, sd, leaf)
        t.factor(pde, tol, kigpu.FP64)
    finally:
        ctx.set_id_method(0)
    rank, active = t.ranks()
    nodes = list(otree.all_nodes())
    orank = np.array([len(n.skel) if n.compressed else -1 for n in nodes], np.int32)
    so, s, ro, r = t.lists()
    boxes = differ = 0
    for i, nd in enumerate(nodes):
        if orank[i] < 0 and rank[i] < 0:
            continue
        boxes += 1
        if rank[i] != orank[i] or list(s[so[i]:so[i + 1]]) != nd.skel:
            differ += 1
    print(f"  {boxes} compressed boxes, {differ} differ from dgeqp3's ID", flush=True)
    assert differ <= max(1, boxes // 100)
    f = b.boundary_values
    mu = t.solve(f)
    mu_ref = ofac.solve(otree, f)
    rel = np.linalg.norm(mu - mu_ref) / np.linalg.norm(mu_ref)
    print(f"  |mu - mu_oracle| / |mu| = {rel:.2e}", flush=True)
    assert rel < 1e-8
    assert LS.backward_error(b, pde, mu) <= 10 * tol
    t.close()


@pytest.mark.gpu
def test_gram_id_laplace_star(ctx):
    print("[gram] Laplace star N=16384 leaf 64 tol 1e-6, FP64 factor with the Gram ID", flush=True)
    compare_with_oracle(ctx, star_problem(16384), K.LAPLACE, 64, 1e-6)


@pytest.mark.gpu
def test_gram_id_below_its_resolution_still_refines(ctx):
    # at id_tol 1e-10 the rank cut sits at remaining norms squared ~1e-20 of
    # |A|^2, below what FP64 resolves in A^T A (the default keeps the
    # pivoted QR there); forced, the Gram ID stops at its resolution floor
    # (~1e-8 relative), and the FP32 factor still refines to 1e-10
    print("[gram] star N=4096 FP32 + refinement at id_tol 1e-10 with the Gram ID forced", flush=True)
    b = star_problem(4096)
    ctx.set_id_method(2)
    try:
        t = gpu_tree(ctx, b)
        t.factor(K.LAPLACE, 1e-10, kigpu.FP32_REFINE)
        mu = t.solve(b.boundary_values, refine_iters=20, rel_residual=1e-10)
        print(f"  {t.last_info}", flush=True)
        assert t.last_info["rel_residual"] <= 1e-10
        assert LS.backward_error(b, K.LAPLACE, mu) <= 1e-9
        t.close()
    finally:
        ctx.set_id_method(0)


@pytest.mark.gpu
def test_gram_id_neumann(ctx):
    print("[gram] Laplace-Neumann star N=4096 leaf 64 tol 1e-6", flush=True)
    b = G.star(4096)
    b.boundary_values = G.bc_point_source(b, (3.0, 3.0))
    compare_with_oracle(ctx, b, K.LAPLACE_NEUMANN, 64, 1e-6)


@pytest.mark.gpu
def test_gram_id_stokes(ctx):
    # (not the circle: its rotational symmetry gives ID columns of exactly
    # equal norms, which neither FP64 route separates reproducibly)
    print("[gram] Stokes star N=2048 leaf 64 tol 1e-6", flush=True)
    b = G.star(2048)
    b.boundary_values = G.bc_stokeslet(b)
    compare_with_oracle(ctx, b, K.STOKES, 64, 1e-6)


@pytest.mark.gpu
def test_fp32_refine_gram_vs_qr(ctx):
    # the FP32 factorization with each ID, refined to 1e-10 against the FP64
    # apply: both converge to the solution of the same FP64 operator
    print("[gram] star N=16384 FP32 + refinement, Gram ID vs pivoted-QR ID", flush=True)
    b = star_problem(16384)
    f = b.boundary_values
    out = {}
    for method in (0, 1):
        ctx.set_id_method(method)
        try:
            t = gpu_tree(ctx, b)
            t.factor(K.LAPLACE, 1e-6, kigpu.FP32_REFINE)
            out[method] = t.solve(f, refine_iters=10, rel_residual=1e-10)
            print(f"  id_method {method}: {t.last_info}", flush=True)
            assert t.last_info["rel_residual"] <= 1e-10
            t.close()
        finally:
            ctx.set_id_method(0)
    rel = np.linalg.norm(out[0] - out[1]) / np.linalg.norm(out[1])
    assert rel < 1e-9
