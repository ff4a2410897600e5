"""The treecode apply of the refinement residual (Laplace double layer, one
or several densities) against the dense kernel of the oracle and the GPU's
direct sum, and the FP32 + refinement solve with it against the same solve
with the direct residual."""
import numpy as np
import pytest

import kigpu
from kigpu import workloads as W
from oracle import geometry as G
from oracle import kernel as K

from helpers import gpu_tree, star_problem

pytestmark = pytest.mark.gpu

# truncation of the series (theta = 1/2, 48 terms) is below 7e-15 of each
# box's sum of |a_j| / distance; 1e-12 of the largest entry of y leaves room
# for the sums' rounding
RTOL = 1e-12


@pytest.mark.parametrize("n,leaf", [(4096, 64), (3000, 32)])
def test_treecode_matches_dense(ctx, n, leaf):
    print(f"[treecode] star {n} leaf {leaf} vs the oracle's dense kernel", flush=True)
    b = G.star(n)
    t = gpu_tree(ctx, b, 1, leaf)
    mu = np.random.default_rng(n).standard_normal(n)
    ref = K.Kernel(1, K.LAPLACE, b)(np.arange(n), np.arange(n)) @ mu
    y = t.apply_treecode(mu)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print(f"  max error / max |y| = {err:.2e}", flush=True)
    assert err <= RTOL
    # a second apply on the same tree (lists kept) with another density
    mu2 = np.cos(np.arange(n) * 0.01)
    ref2 = K.Kernel(1, K.LAPLACE, b)(np.arange(n), np.arange(n)) @ mu2
    assert np.abs(t.apply_treecode(mu2) - ref2).max() <= RTOL * np.abs(ref2).max()
    t.close()


@pytest.mark.parametrize("geom", ["star", "random_star", "circle"])
def test_treecode_matches_direct_large(ctx, geom):
    n = 65536
    print(f"[treecode] {geom} {n} vs the direct GPU sum", flush=True)
    if geom == "star":
        pts, nrm, w, kap = W.star(n)
    elif geom == "random_star":
        pts, nrm, w, kap = W.random_star(n, 5)
    else:
        b = G.circle(n)
        pts, nrm, w, kap = b.points, b.normals, b.weights, b.curvatures
    t = ctx.build_tree(pts, nrm, w, kap, 1, 64)
    mu = np.random.default_rng(3).standard_normal(n)
    ref = t.apply(mu, kigpu.LAPLACE)
    y = t.apply_treecode(mu)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print(f"  max error / max |y| = {err:.2e}", flush=True)
    assert err <= RTOL
    t.close()


@pytest.mark.parametrize("nrhs", [3, 17])
def test_treecode_many_rhs(ctx, nrhs):
    # several densities in one apply (one grid row per column, shared lists):
    # each column against the direct sum and against the one-density apply
    n = 20000
    print(f"[treecode] random_star {n}, {nrhs} densities", flush=True)
    pts, nrm, w, kap = W.random_star(n, 9)
    t = ctx.build_tree(pts, nrm, w, kap, 1, 64)
    mu = np.random.default_rng(nrhs).standard_normal((n, nrhs))
    ref = t.apply(mu, kigpu.LAPLACE)
    y = t.apply_treecode(mu)
    assert y.shape == ref.shape
    err = (np.abs(y - ref).max(axis=0) / np.abs(ref).max(axis=0)).max()
    print(f"  worst column: max error / max |y| = {err:.2e}", flush=True)
    assert err <= RTOL
    for j in (0, nrhs - 1):
        assert np.array_equal(y[:, j], t.apply_treecode(np.ascontiguousarray(mu[:, j])))
    t.close()


def test_treecode_forest(ctx):
    # block-diagonal over instances of uneven sizes
    print("[treecode] forest of 4 instances", flush=True)
    sizes = [300, 512, 1100, 2049]
    geoms = [W.random_star(n, 70 + i) for i, n in enumerate(sizes)]
    f = ctx.build_batch(geoms, 1, 64)
    pt_off = f.instances()["point_offsets"]
    mu = np.random.default_rng(8).standard_normal(int(pt_off[-1]))
    y = f.apply_treecode(mu)
    for i, n in enumerate(sizes):
        b = G.random_star(n, 70 + i)
        a, e = pt_off[i], pt_off[i + 1]
        ref = K.Kernel(1, K.LAPLACE, b)(np.arange(n), np.arange(n)) @ mu[a:e]
        assert np.abs(y[a:e] - ref).max() <= RTOL * np.abs(ref).max()
    f.close()


def test_refinement_treecode_vs_direct(ctx):
    # the FP32 factor refined to 1e-10 with the treecode residual and with the
    # direct one: both converge, the treecode's density also has a direct-sum
    # residual below 1e-10, and the two densities agree to the refinement's
    # tolerance
    print("[treecode] FP32 + refinement, treecode vs direct residual", flush=True)
    b = star_problem(16384)
    t = gpu_tree(ctx, b, 1, 64)
    t.factor(kigpu.LAPLACE, 1e-6, kigpu.FP32_REFINE)
    mus = []
    for mode in (0, 1):
        ctx.set_refine_apply(mode)
        mus.append(t.solve(b.boundary_values, refine_iters=10, rel_residual=1e-10))
        print(f"  mode {mode}: {t.last_info}", flush=True)
        assert t.last_info["rel_residual"] <= 1e-10
        tm = ctx.timings()
        assert (tm["tc_launches"] > 0) == (mode == 0)
        assert (tm["apply_launches"] > 0) == (mode == 1)
    ctx.set_refine_apply(0)
    f = b.boundary_values
    res = np.linalg.norm(f - t.apply(mus[0], kigpu.LAPLACE)) / np.linalg.norm(f)
    d = np.abs(mus[0] - mus[1]).max() / np.abs(mus[1]).max()
    print(f"  direct residual of the treecode-refined density {res:.2e}; "
          f"densities differ by {d:.2e} (relative)", flush=True)
    assert res <= 1e-10
    assert d <= 1e-9
    t.close()


def test_refinement_treecode_many_rhs(ctx):
    # eight right-hand sides refined with the treecode residual and with the
    # direct one; every column converges and the two agree
    print("[treecode] FP32 + refinement, 8 right-hand sides", flush=True)
    b = star_problem(8192)
    t = gpu_tree(ctx, b, 1, 64)
    t.factor(kigpu.LAPLACE, 1e-6, kigpu.FP32_REFINE)
    rng = np.random.default_rng(21)
    F = np.column_stack([b.boundary_values] + [rng.standard_normal(b.num_points)
                                                for _ in range(7)])
    mus = []
    for mode in (0, 1):
        ctx.set_refine_apply(mode)
        mus.append(t.solve(F, refine_iters=10, rel_residual=1e-10))
        print(f"  mode {mode}: {t.last_info}", flush=True)
        assert t.last_info["rel_residual"] <= 1e-10
        tm = ctx.timings()
        assert (tm["tc_launches"] > 0) == (mode == 0)
    ctx.set_refine_apply(0)
    R = F - t.apply(mus[0], kigpu.LAPLACE)
    res = (np.linalg.norm(R, axis=0) / np.linalg.norm(F, axis=0)).max()
    d = (np.abs(mus[0] - mus[1]).max(axis=0) / np.abs(mus[1]).max(axis=0)).max()
    print(f"  worst direct residual {res:.2e}; densities differ by {d:.2e}", flush=True)
    assert res <= 1e-10
    assert d <= 1e-9
    t.close()
