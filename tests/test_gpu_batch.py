"""Batched uneven instances: one forest factored with one launch per level
over the boxes of every instance, against one tree per instance.

Bars: every instance's quadtree (node order, parents, children, corners,
leaf flags, DOF lists, neighbor lists) identical to its single-tree build;
per-box ranks identical and mu equal to the single-tree FP64 solve (the
per-box kernels and their reductions are the same, so equal to the last
bit is expected; the bar allows 1e-14); FP32 + refinement: every instance's
residual against its own FP64 kernel below the requested 1e-10."""
import numpy as np
import pytest

import kigpu
from kigpu import workloads as W
from oracle import geometry as G
from oracle import kernel as K


def instance_geoms(count, seed, lo=1024, hi=4096):
    sizes = W.batch_sizes(count, seed, lo, hi)
    return [W.random_star(int(n), seed * 1000 + i) for i, n in enumerate(sizes)]


def forest_instances(nodes):
    """node -> instance by following parents (roots are nodes 0..ninst-1)."""
    par = nodes["parent"]
    inst = np.full(par.size, -1, np.int64)
    lo = nodes["level_offsets"]
    inst[: lo[1]] = np.arange(lo[1])
    for x in range(lo[1], par.size):
        inst[x] = inst[par[x]]
    return inst


def split_forest(t):
    nodes = t.export_nodes()
    inst = forest_instances(nodes)
    lo = nodes["level_offsets"]
    level = np.zeros(inst.size, np.int64)
    for l in range(lo.size - 1):
        level[lo[l]:lo[l + 1]] = l
    return nodes, inst, level


@pytest.mark.gpu
def test_forest_trees_match_single_trees(ctx):
    print("[batch] forest of 6 random stars vs single trees", flush=True)
    geoms = instance_geoms(6, 11)
    f = ctx.build_batch(geoms, 1, 64)
    nodes, inst, level = split_forest(f)
    boff, boxes = f.export_boxes()
    noff, nbrs = f.export_neighbors()
    pt_off = f.instances()["point_offsets"]
    for i, g in enumerate(geoms):
        s = ctx.build_tree(*g, 1, 64)
        sn = s.export_nodes()
        mine = np.nonzero(inst == i)[0]           # forest ids, level by level in order
        # same node count per level and node order within levels
        L = sn["level_offsets"].size - 1
        assert np.array_equal(np.bincount(level[mine], minlength=L)[:L],
                              np.diff(sn["level_offsets"]))
        order = np.concatenate([mine[level[mine] == l] for l in range(L)])
        to_local = {int(x): k for k, x in enumerate(order)}
        mapn = lambda a: np.array([to_local[int(v)] if v >= 0 else -1 for v in a])
        assert np.array_equal(mapn(nodes["parent"][order]), sn["parent"])
        assert np.array_equal(mapn(nodes["children"][order].ravel()), sn["children"].ravel())
        assert np.array_equal(nodes["is_leaf"][order], sn["is_leaf"])
        assert np.array_equal(nodes["corners"][order], sn["corners"])
        so, sb = s.export_boxes()
        sno, snb = s.export_neighbors()
        for k, x in enumerate(order):
            got = boxes[boff[x]:boff[x + 1]] - pt_off[i]
            assert np.array_equal(got, sb[so[k]:so[k + 1]])
            assert np.array_equal(mapn(nbrs[noff[x]:noff[x + 1]]), snb[sno[k]:sno[k + 1]])
        s.close()
    f.close()


@pytest.mark.gpu
def test_forest_factor_solve_matches_single(ctx):
    print("[batch] FP64 factor/solve of 6 instances vs single trees", flush=True)
    geoms = instance_geoms(6, 12)
    f = ctx.build_batch(geoms, 1, 64)
    f.factor(kigpu.LAPLACE, 1e-6, kigpu.FP64)
    nodes, inst, level = split_forest(f)
    rank, _ = f.ranks()
    rhs = [np.stack([W.point_source_bc(g[0], tuple(s)) for s in W.point_sources(3, 5)], 1)
           for g in geoms]
    X = f.solve(np.concatenate(rhs, 0))
    pt_off = f.instances()["point_offsets"]
    worst = 0.0
    for i, g in enumerate(geoms):
        s = ctx.build_tree(*g, 1, 64)
        s.factor(kigpu.LAPLACE, 1e-6, kigpu.FP64)
        srank, _ = s.ranks()
        mine = np.nonzero(inst == i)[0]
        L = int(level[mine].max()) + 1
        order = np.concatenate([mine[level[mine] == l] for l in range(L)])
        assert np.array_equal(rank[order], srank)
        ref = s.solve(rhs[i])
        got = X[pt_off[i]:pt_off[i + 1]]
        d = np.abs(got - ref).max() / np.abs(ref).max()
        worst = max(worst, d)
        assert d <= 1e-14
        s.close()
    print(f"  worst instance difference {worst:.3e}", flush=True)
    f.close()


@pytest.mark.gpu
def test_forest_apply_one_density(ctx):
    # block-diagonal apply over instances smaller than, equal to and larger
    # than one 512-node block, against each instance's dense matrix
    print("[batch] one-density apply over 4 instances", flush=True)
    sizes = [300, 512, 1100, 2049]
    geoms = [W.random_star(n, 70 + i) for i, n in enumerate(sizes)]
    f = ctx.build_batch(geoms, 1, 64)
    pt_off = f.instances()["point_offsets"]
    mu = np.random.default_rng(8).standard_normal(int(pt_off[-1]))
    y = f.apply(mu, kigpu.LAPLACE)
    for i, n in enumerate(sizes):
        b = G.random_star(n, 70 + i)
        assert np.array_equal(b.points, geoms[i][0])
        a, e = pt_off[i], pt_off[i + 1]
        ref = K.Kernel(1, K.LAPLACE, b)(np.arange(n), np.arange(n)) @ mu[a:e]
        assert np.allclose(y[a:e], ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    f.close()


@pytest.mark.gpu
def test_forest_apply_many_rhs(ctx):
    # the matrix-core apply over instances of uneven sizes, 24 right-hand sides
    print("[batch] 24-column apply over 4 instances", flush=True)
    sizes = [300, 512, 1100, 2049]
    geoms = [W.random_star(n, 70 + i) for i, n in enumerate(sizes)]
    f = ctx.build_batch(geoms, 1, 64)
    pt_off = f.instances()["point_offsets"]
    mu = np.random.default_rng(9).standard_normal((int(pt_off[-1]), 24))
    y = f.apply(mu, kigpu.LAPLACE)
    for i, n in enumerate(sizes):
        b = G.random_star(n, 70 + i)
        a, e = pt_off[i], pt_off[i + 1]
        ref = K.Kernel(1, K.LAPLACE, b)(np.arange(n), np.arange(n)) @ mu[a:e]
        assert np.allclose(y[a:e], ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    f.close()


@pytest.mark.gpu
def test_forest_fp32_refined(ctx):
    print("[batch] FP32 + refinement, 8 instances", flush=True)
    geoms = instance_geoms(8, 13)
    f = ctx.build_batch(geoms, 1, 64)
    f.factor(kigpu.LAPLACE, 1e-6, kigpu.FP32_REFINE)
    rhs = np.concatenate([W.point_source_bc(g[0]) for g in geoms])
    mu = f.solve(rhs, refine_iters=10, rel_residual=1e-10)
    assert f.last_info["rel_residual"] <= 1e-10
    # the forest's apply is block diagonal: each block equals the instance's own
    y = f.apply(mu, kigpu.LAPLACE)
    pt_off = f.instances()["point_offsets"]
    for i, g in enumerate(geoms):
        s = ctx.build_tree(*g, 1, 64)
        a, b = pt_off[i], pt_off[i + 1]
        yi = s.apply(mu[a:b], kigpu.LAPLACE)
        assert np.allclose(y[a:b], yi, rtol=0, atol=1e-13 * np.abs(yi).max())
        fi = rhs[a:b]
        assert np.linalg.norm(yi - fi) <= 1e-10 * np.linalg.norm(fi)
        s.close()
    f.close()


@pytest.mark.gpu
def test_batch_256_uneven(ctx):
    # 256 stars r = 1 + sum_{k=2..6} a_k cos(k t + phi_k), N ~ U{1024..16384}
    # rounded to 64 (about 2.2M DOFs), FP32 + refinement to 1e-10
    print("[batch] 256 uneven instances, FP32 + refinement", flush=True)
    sizes = W.batch_sizes(256, 4)
    geoms = [W.random_star(int(n), 4000 + i) for i, n in enumerate(sizes)]
    f = ctx.build_batch(geoms, 1, 64)
    f.factor(kigpu.LAPLACE, 1e-6, kigpu.FP32_REFINE)
    rhs = np.concatenate([W.point_source_bc(g[0]) for g in geoms])
    mu = f.solve(rhs, refine_iters=20, rel_residual=1e-10)
    print(f"  {int(sizes.sum())} DOFs, refinement {f.last_info}", flush=True)
    assert f.last_info["rel_residual"] <= 1e-10
    pt_off = f.instances()["point_offsets"]
    for i in (0, 101, 255):
        s = ctx.build_tree(*geoms[i], 1, 64)
        s.factor(kigpu.LAPLACE, 1e-6, kigpu.FP32_REFINE)
        ref = s.solve(rhs[pt_off[i]:pt_off[i + 1]], refine_iters=20, rel_residual=1e-10)
        got = mu[pt_off[i]:pt_off[i + 1]]
        assert np.linalg.norm(got - ref) <= 1e-9 * np.linalg.norm(ref)
        s.close()
    f.close()
