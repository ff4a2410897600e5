"""Pins for the oracle quadtree (sequential insertion, quadtree.cpp:72-285).

* node set, levels and DOF lists equal a brute-force recursive partition;
* node order inside each level equals the level-synchronous rule
  (splitting nodes sorted by (effective split point index, child-slot path),
  children in slot order) -- a different algorithm deriving the same order;
* same-level neighbor sets equal the geometric 8-adjacency of equal boxes;
* leaves hold at most leaf_dofs DOFs, in ascending order, and partition the DOFs.
"""
import numpy as np
import pytest

from oracle import geometry as G
from oracle.tree import QuadTree, _quadrant

CASES = [
    ("star", lambda: G.star(4096), 1, 64),
    ("circle", lambda: G.circle(1024), 1, 32),
    ("annulus", lambda: G.annulus(1024), 1, 32),
    ("holes_stokes", lambda: G.multi_hole_disc(2048, 256, 8, 7), 2, 64),
    ("rstar", lambda: G.random_star(3008, 11), 1, 64),
]


def brute_force(points, sd, leaf):
    xy = np.asarray(points).reshape(-1, 2)
    lo, hi = float(np.min(points)) - 0.01, float(np.max(points)) + 0.01
    out = {}

    def rec(corners, side, level, path, pts):
        dofs = [sd * p + c for p in pts for c in range(sd)]
        out[(level, path)] = (dofs, list(corners), side)
        if len(dofs) <= leaf:
            return
        midx = ((corners[6] - corners[0]) / 2.0) + corners[0]
        midy = ((corners[3] - corners[1]) / 2.0) + corners[1]
        groups = [[], [], [], []]
        for p in pts:
            groups[_quadrant(xy[p, 0], xy[p, 1], midx, midy)].append(p)
        c = corners
        kids = [
            [c[0], c[1], c[2], midy, midx, midy, midx, c[7]],
            [c[0], midy, c[2], c[3], midx, c[5], midx, midy],
            [midx, midy, midx, c[3], c[4], c[5], c[6], midy],
            [midx, c[1], midx, midy, c[4], midy, c[6], c[7]],
        ]
        for s in range(4):
            rec(kids[s], side / 2.0, level + 1, path * 4 + s, groups[s])

    rec([lo, lo, lo, hi, hi, hi, hi, lo], hi - lo, 0, 0, list(range(xy.shape[0])))
    return out


def level_sync_order(tree):
    """Order of every level from the level-synchronous rule."""
    sd, leaf = tree.sd, tree.leaf_dofs
    orders = [[tree.root]]
    eff = {}
    while True:
        split = []
        for nd in orders[-1]:
            if len(nd.original_box) > leaf:
                t = nd.original_box[(leaf // sd) * sd] // sd
                par = nd.parent
                te = max(t, eff[id(par)]) if par is not None else t
                eff[id(nd)] = te
                split.append((te, nd.path, nd))
        if not split:
            return orders
        split.sort(key=lambda e: (e[0], e[1]))
        orders.append([ch for _, _, nd in split for ch in nd.children])


@pytest.mark.parametrize("name,make,sd,leaf", CASES, ids=[c[0] for c in CASES])
def test_tree_against_brute_force(name, make, sd, leaf):
    b = make()
    tree = QuadTree(b.points, sd, leaf)
    bf = brute_force(b.points, sd, leaf)
    got = {(n.level, n.path): (n.original_box, n.corners, n.side) for n in tree.all_nodes()}
    assert got.keys() == bf.keys()
    for key, (dofs, corners, side) in bf.items():
        g = got[key]
        assert g[0] == dofs
        assert g[1] == corners and g[2] == side
    all_leaf = []
    for n in tree.all_nodes():
        if n.is_leaf:
            assert len(n.original_box) <= leaf
            assert n.original_box == sorted(n.original_box)
            all_leaf += n.original_box
    assert sorted(all_leaf) == list(range(sd * b.num_points))


@pytest.mark.parametrize("name,make,sd,leaf", CASES, ids=[c[0] for c in CASES])
def test_node_order_rule(name, make, sd, leaf):
    b = make()
    tree = QuadTree(b.points, sd, leaf)
    orders = level_sync_order(tree)
    assert len(orders) == len(tree.levels)
    for lvl, order in zip(tree.levels, orders):
        assert [id(n) for n in lvl] == [id(n) for n in order]


@pytest.mark.parametrize("name,make,sd,leaf", CASES, ids=[c[0] for c in CASES])
def test_same_level_neighbors_geometric(name, make, sd, leaf):
    b = make()
    tree = QuadTree(b.points, sd, leaf)
    for lvl in tree.levels:
        for a in lvl:
            got = {id(n) for n in a.neighbors if n.level == a.level}
            want = set()
            for c in lvl:
                if c is a:
                    continue
                dx = abs(a.corners[0] - c.corners[0])
                dy = abs(a.corners[1] - c.corners[1])
                if dx <= a.side * 1.000001 and dy <= a.side * 1.000001:
                    want.add(id(c))
            assert got == want


def _fixture():
    import json
    import os
    p = os.path.join(os.path.dirname(__file__), "golden", "tree_trace_4pts.json")
    with open(p) as fh:
        return json.load(fh)


def _expected_boxes(fx, sd):
    return [[sd * p + c for p in pts for c in range(sd)] for pts in fx["original_box_points"]]


@pytest.mark.parametrize("case", [0, 1, 2])
def test_hand_traced_tree(case):
    # node order (subdivision events, not slot paths), DOF lists and neighbor
    # lists -- including the coarse leaves that get_descendent_neighbors
    # pushes and the contacts of the `y < top && y < bottom` test
    # (quadtree.cpp:143) -- against the trace in tests/golden
    fx = _fixture()
    sd, leaf = fx["cases"][case]["sd"], fx["cases"][case]["leaf_dofs"]
    ex = QuadTree(np.array(fx["points"]), sd, leaf).export()
    assert ex["level_offsets"].tolist() == fx["level_offsets"]
    assert ex["parent"].tolist() == fx["parent"]
    assert ex["children"].tolist() == fx["children"]
    assert ex["is_leaf"].tolist() == fx["is_leaf"]
    boxes = _expected_boxes(fx, sd)
    off, dofs = ex["box_offsets"], ex["box_dofs"]
    assert [dofs[off[i]:off[i + 1]].tolist() for i in range(len(boxes))] == boxes
    noff, nbrs = ex["nbr_offsets"], ex["nbrs"]
    assert [nbrs[noff[i]:noff[i + 1]].tolist() for i in range(len(boxes))] == fx["neighbors"]


@pytest.mark.parametrize("name,make,sd,leaf", CASES, ids=[c[0] for c in CASES])
def test_child_boxes_contain_their_points(name, make, sd, leaf):
    # child slots checked by containment in the child's own corners
    # (half-open: a point on a midline belongs to the right / upper child,
    # quadtree.cpp:177-185), and each box is the union of its children's
    b = make()
    tree = QuadTree(b.points, sd, leaf)
    xy = b.points.reshape(-1, 2)
    for n in tree.all_nodes():
        c = n.corners
        for d in n.original_box[::sd]:
            x, y = xy[d // sd]
            assert c[0] <= x and c[1] <= y
            assert (x < c[4] or n.parent is None) and (y < c[3] or n.parent is None)
        if not n.is_leaf:
            kids = sorted(d for ch in n.children for d in ch.original_box)
            assert kids == sorted(n.original_box)
