"""GPU quadtree vs the sequential-insertion oracle: bit-exact node order,
corners, DOF lists and neighbor lists (quadtree.cpp:27-285)."""
import types

import numpy as np
import pytest

from oracle import geometry as G
from oracle.tree import QuadTree

from helpers import gpu_tree

def clustered():
    # 48 points within 1e-7 of (0.1, 0.2) inside a unit circle of 64: the
    # tree runs deeper than 20 levels (the pipelined build's sort limit),
    # so the build restarts with a wait per level
    b = G.circle(64)
    i = np.arange(48)
    extra = np.stack([0.1 + 2e-9 * i, 0.2 + 1.5e-9 * (i % 7)], 1)
    pts = np.concatenate([b.points.reshape(-1, 2), extra]).reshape(-1)
    n = pts.size // 2
    return types.SimpleNamespace(points=pts, normals=np.tile([1.0, 0.0], n),
                                 weights=np.full(n, 1.0 / n), curvatures=np.zeros(n))


CASES = [
    ("star4096_leaf64", lambda: G.star(4096), 1, 64),
    ("clustered_deep", clustered, 1, 32),
    ("circle1024_leaf32", lambda: G.circle(1024), 1, 32),
    ("annulus1024_leaf32", lambda: G.annulus(1024), 1, 32),
    ("circle1024_stokes_leaf32", lambda: G.circle(1024), 2, 32),
    ("holes_stokes", lambda: G.multi_hole_disc(2048, 256, 8, 7), 2, 64),
    ("rstar_leaf64", lambda: G.random_star(3008, 11), 1, 64),
    ("tiny3", lambda: G.circle(3), 1, 32),
    ("star65536_leaf64", lambda: G.star(65536), 1, 64),
]


@pytest.mark.gpu
@pytest.mark.parametrize("name,make,sd,leaf", CASES, ids=[c[0] for c in CASES])
def test_tree_bit_exact(ctx, name, make, sd, leaf):
    print(f"[tree] {name}", flush=True)
    b = make()
    ref = QuadTree(b.points, sd, leaf).export()
    t = gpu_tree(ctx, b, sd, leaf)
    got = t.export_nodes()
    assert np.array_equal(got["level_offsets"], ref["level_offsets"])
    assert np.array_equal(got["parent"], ref["parent"])
    assert np.array_equal(got["children"], ref["children"])
    assert np.array_equal(got["is_leaf"], ref["is_leaf"])
    # corners and sides are compared bit for bit
    assert np.array_equal(got["corners"].view(np.int64), ref["corners"].view(np.int64))
    assert np.array_equal(got["side"].view(np.int64), ref["side"].view(np.int64))
    off, dofs = t.export_boxes()
    assert np.array_equal(off, ref["box_offsets"])
    assert np.array_equal(dofs, ref["box_dofs"])
    noff, nbrs = t.export_neighbors()
    assert np.array_equal(noff, ref["nbr_offsets"])
    assert np.array_equal(nbrs, ref["nbrs"])
    info = t.info()
    assert info["min"] == float(np.min(b.points)) and info["max"] == float(np.max(b.points))


@pytest.mark.gpu
@pytest.mark.parametrize("case", [0, 1, 2])
def test_tree_hand_traced(ctx, case):
    # the four-point tree traced by hand through quadtree.cpp:27-285
    # (tests/golden/tree_trace_4pts.md)
    import json
    import os
    print(f"[tree] hand-traced fixture case {case}", flush=True)
    with open(os.path.join(os.path.dirname(__file__), "golden", "tree_trace_4pts.json")) as fh:
        fx = json.load(fh)
    sd, leaf = fx["cases"][case]["sd"], fx["cases"][case]["leaf_dofs"]
    pts = np.array(fx["points"])
    n = pts.size // 2
    t = ctx.build_tree(pts, np.tile([1.0, 0.0], n), np.full(n, 1.0 / n), np.zeros(n), sd, leaf)
    got = t.export_nodes()
    assert got["level_offsets"].tolist() == fx["level_offsets"]
    assert got["parent"].tolist() == fx["parent"]
    assert got["children"].tolist() == fx["children"]
    assert got["is_leaf"].tolist() == fx["is_leaf"]
    off, dofs = t.export_boxes()
    boxes = [[sd * p + c for p in b for c in range(sd)] for b in fx["original_box_points"]]
    assert [dofs[off[i]:off[i + 1]].tolist() for i in range(len(boxes))] == boxes
    noff, nbrs = t.export_neighbors()
    assert [nbrs[noff[i]:noff[i + 1]].tolist() for i in range(len(boxes))] == fx["neighbors"]
