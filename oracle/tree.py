"""Adaptive quadtree over boundary nodes: sequential-insertion restatement.

TEST INFRASTRUCTURE ONLY (see the package docstring).

Restates ``kern_interp/quadtree/quadtree.cpp`` (construction ``:72-114``,
point insertion ``:164-197``, subdivision ``:203-285``, neighbor lists
``:27-69`` and ``:119-161``, active/near DOF lists ``:656-720``, the
incremental edits ``perturb``, ``mark_neighbors_and_parents`` and
``consolidate_node`` ``:288-560``) and the per-node DOF bookkeeping of
``quadtree.h:14-61``.  The leaf capacity, fixed
at 32 DOFs by ``quadtree.h:10``, is a parameter here.

The insertion is simulated point by point, exactly as the reference does, so
node order inside each level is the order of subdivision events.  The GPU
tree uses a different, level-synchronous algorithm; agreement of the two is
one of the pins.  Other pins (``tests/test_oracle_tree.py``): a brute-force
recursive partition gives the same node set and DOF lists, every leaf holds
at most ``leaf_dofs`` DOFs in ascending order, and neighbor sets match a
geometric adjacency test.  The edits are pinned by two perturbations of the
four-point tree traced by hand (``tests/golden/perturb_trace_4pts.md``) and
compared with the drop-in's C++ host code (``tests/test_perturb_host.py``).
"""
from __future__ import annotations

import math

import numpy as np

# corner layout of a node: [BL.x, BL.y, TL.x, TL.y, TR.x, TR.y, BR.x, BR.y]
SQRT2 = math.sqrt(2)


class Node:
    __slots__ = ("level", "index", "is_leaf", "parent", "children", "corners",
                 "side", "neighbors", "original_box", "active_box", "skel",
                 "redundant", "near", "skelnear", "permutation", "compressed",
                 "T", "L", "U", "X_lu", "X_piv", "schur", "path", "split_time",
                 "id", "m_rows", "dofs_below")

    def __init__(self, level, corners, side, parent):
        self.level = level
        self.index = -1
        self.is_leaf = True
        self.parent = parent
        self.children = [None, None, None, None]
        self.corners = list(corners)
        self.side = side
        self.neighbors = []
        self.original_box = []
        self.active_box = []
        self.skel = []
        self.redundant = []
        self.near = []
        self.skelnear = []
        self.permutation = []
        self.compressed = False
        self.T = self.L = self.U = self.X_lu = self.X_piv = self.schur = None
        self.path = 0
        self.split_time = -1
        self.id = -1
        self.m_rows = 0
        self.dofs_below = 0

    def mid(self):
        c = self.corners
        return ((c[6] - c[0]) / 2.0) + c[0], ((c[3] - c[1]) / 2.0) + c[1]


def _quadrant(x, y, midx, midy):
    """Child slot for a point: BL=0, TL=1, TR=2, BR=3, strict '<' (quadtree.cpp:177-185)."""
    if x < midx:
        return 0 if y < midy else 1
    return 3 if y < midy else 2


class QuadTree:
    """Sequential restatement of ``QuadTree::initialize_tree`` (quadtree.cpp:72-114)."""

    def __init__(self, points, sd: int, leaf_dofs: int = 32):
        pts = np.asarray(points, dtype=np.float64).reshape(-1)
        assert pts.size > 0, "number of boundary points cannot be 0"
        self.points = pts
        self.sd = sd
        self.leaf_dofs = leaf_dofs
        # root box: min/max pooled over x and y (quadtree.cpp:80-88)
        self.min = float(pts.min())
        self.max = float(pts.max())
        lo, hi = self.min - 0.01, self.max + 0.01
        self.root = Node(0, [lo, lo, lo, hi, hi, hi, hi, lo], hi - lo, None)
        self.levels = [[self.root]]
        self._time = 0
        xy = pts.reshape(-1, 2)
        self._x = xy[:, 0].tolist()
        self._y = xy[:, 1].tolist()
        for i in range(xy.shape[0]):
            self._time = i
            self._insert(i)
        self._number()
        self._neighbor_lists()
        # factor-time state
        self.allskel = []
        self.allskel_lu = None
        self.allskel_piv = None
        self.U = None
        self.Psi = None
        self.S_lu = None
        self.S_piv = None

    # -- construction ------------------------------------------------------
    def _insert(self, i):
        """``recursive_add`` (quadtree.cpp:164-197), iteratively."""
        x, y = self._x[i], self._y[i]
        node = self.root
        dofs = [self.sd * i + c for c in range(self.sd)]
        while True:
            node.original_box.extend(dofs)
            if node.is_leaf:
                if len(node.original_box) > self.leaf_dofs:
                    self._subdivide(node)
                return
            midx, midy = node.mid()
            node = node.children[_quadrant(x, y, midx, midy)]

    def _subdivide(self, node):
        """``node_subdivide`` (quadtree.cpp:203-285): children appended to the next level."""
        node.is_leaf = False
        node.split_time = self._time
        midx, midy = node.mid()
        c = node.corners
        kids = []
        # BL, TL, TR, BR child corners from the parent's corners and midpoint
        for slot in range(4):
            k = list(c)
            if slot == 0:
                k[3], k[4], k[5], k[6] = midy, midx, midy, midx
            elif slot == 1:
                k[1], k[4], k[6], k[7] = midy, midx, midx, midy
            elif slot == 2:
                k[0], k[1], k[2], k[7] = midx, midy, midx, midy
            else:
                k[0], k[2], k[3], k[5] = midx, midx, midy, midy
            ch = Node(node.level + 1, k, node.side / 2.0, node)
            ch.path = node.path * 4 + slot
            kids.append(ch)
        node.children = kids
        if len(self.levels) < node.level + 2:
            self.levels.append([])
        self.levels[node.level + 1].extend(kids)
        sd = self.sd
        for j in range(0, len(node.original_box), sd):
            dof = node.original_box[j]
            p = dof // sd
            q = _quadrant(self._x[p], self._y[p], midx, midy)
            kids[q].original_box.extend(range(dof, dof + sd))
        for ch in kids:
            if len(ch.original_box) > self.leaf_dofs:
                self._subdivide(ch)

    def _number(self):
        nid = 0
        for lvl in self.levels:
            for i, nd in enumerate(lvl):
                nd.index = i
                nd.id = nid
                nid += 1
        self.num_nodes = nid

    def _neighbor_lists(self):
        """``compute_neighbor_lists`` (quadtree.cpp:27-69)."""
        for lvl in self.levels:
            for a in lvl:
                par = a.parent
                if par is not None:
                    for sib in par.children:
                        if sib is not a:
                            a.neighbors.append(sib)
                    thr = a.side * SQRT2 + 1e-5
                    for pn in par.neighbors:
                        for cz in pn.children:
                            if cz is None or cz.level != a.level:
                                continue
                            dx = a.corners[0] - cz.corners[0]
                            dy = a.corners[1] - cz.corners[1]
                            if math.sqrt(dx * dx + dy * dy) < thr:
                                a.neighbors.append(cz)
                if a.is_leaf:
                    n = 0
                    while n < len(a.neighbors):
                        nb = a.neighbors[n]
                        n += 1
                        if nb.level != a.level:
                            continue
                        for ch in nb.children:
                            if ch is not None:
                                self._descendent_neighbors(a, ch)

    def _descendent_neighbors(self, big, small):
        """``get_descendent_neighbors`` (quadtree.cpp:119-161), including the
        left/right wall test ``y < top && y < bottom`` of ``:143`` as written."""
        top, bottom = big.corners[3], big.corners[1]
        left, right = big.corners[0], big.corners[4]
        for i in range(0, 8, 2):
            x, y = small.corners[i], small.corners[i + 1]
            hit = False
            if left < x < right:
                if abs(y - top) < 1e-14 or abs(y - bottom) < 1e-14:
                    hit = True
            if not hit and y < top and y < bottom:
                if abs(x - left) < 1e-14 or abs(x - right) < 1e-14:
                    hit = True
            if hit:
                big.neighbors.append(small)
                small.neighbors.append(big)
                break
        for ch in small.children:
            if ch is not None:
                self._descendent_neighbors(big, ch)

    # -- DOF bookkeeping ---------------------------------------------------
    @staticmethod
    def refresh_active(node):
        """``remove_inactive_dofs_at_box`` (quadtree.cpp:694-720)."""
        node.skel = []
        node.redundant = []
        if node.is_leaf:
            node.active_box = list(node.original_box)
        else:
            ab = []
            for ch in node.children:
                ab.extend(ch.skel if ch.compressed else ch.active_box)
            node.active_box = ab

    def refresh_level(self, level):
        """``remove_inactive_dofs_at_level`` (quadtree.cpp:664-691)."""
        nodes = self.levels[level]
        for nd in nodes:
            if not nd.compressed:
                self.refresh_active(nd)
        for a in nodes:
            near = []
            for nb in a.neighbors:
                if nb.level > a.level:
                    continue
                near.extend(nb.original_box if nb.is_leaf else nb.active_box)
            a.near = near

    def refresh_all(self):
        """``remove_inactive_dofs_at_all_boxes`` (quadtree.cpp:656-661)."""
        for level in range(len(self.levels) - 1, -1, -1):
            self.refresh_level(level)

    # -- incremental edits (quadtree.cpp:288-560) -----------------------------
    @staticmethod
    def mark_neighbors_and_parents(node):
        """``mark_neighbors_and_parents`` (quadtree.cpp:288-313): the box, its
        neighbors, and every ancestor with its neighbors lose their
        compression and factor blocks."""
        while node is not None:
            for n in [node] + node.neighbors:
                n.compressed = False
                n.active_box = []
                n.skel = []
                n.skelnear = []
                n.redundant = []
                n.permutation = []
                n.T = n.U = n.L = None
            node = node.parent

    def consolidate_node(self, node):
        """``consolidate_node`` (quadtree.cpp:316-356): the box becomes a leaf
        holding its leaves' DOFs in breadth-first order; its descendants leave
        their levels."""
        node.original_box = []
        queue, gone = [node], []
        i = 0
        while i < len(queue):
            cur = queue[i]
            i += 1
            if cur.is_leaf:
                node.original_box.extend(cur.original_box)
            else:
                queue.extend(cur.children)
            if cur is not node:
                gone.append(cur)
        for g in gone:
            lvl = self.levels[g.level]
            for k, m in enumerate(lvl):
                if m is g:
                    del lvl[k]
                    break
        node.children = [None, None, None, None]
        node.is_leaf = True

    def _leaf_of(self, x, y):
        cur = self.root
        while not cur.is_leaf:
            midx, midy = cur.mid()
            cur = cur.children[_quadrant(x, y, midx, midy)]
        return cur

    def perturb(self, new_points):
        """``QuadTree::perturb`` (quadtree.cpp:359-560).  Points are matched by
        exact coordinates (a repeated point keeps its last new index, as the
        reference's hash map does); the DOF lists are renumbered (a leaf's
        original box, every box's active, skeleton, redundant and skelnear
        lists; a non-leaf's original box is left as it was); added points go
        to the leaves holding them, deleted points mark the first box on
        their path holding fewer than ``leaf_dofs`` DOFs, which is then
        merged into one leaf; overfull leaves subdivide; the neighbor lists
        are computed afresh.  Empty trailing levels are kept, as there."""
        sd = self.sd
        old = self.points.reshape(-1, 2)
        new = np.asarray(new_points, dtype=np.float64).reshape(-1, 2)
        new_index = {}
        for i in range(new.shape[0]):
            new_index[(float(new[i, 0]), float(new[i, 1]))] = i
        found = [False] * new.shape[0]
        old_to_new, deletions = {}, []
        for i in range(old.shape[0]):
            j = new_index.get((float(old[i, 0]), float(old[i, 1])))
            if j is None:
                deletions.append(i)
            else:
                old_to_new[i] = j
                found[j] = True
        additions = [i for i in range(new.shape[0]) if not found[i]]

        def remap(v):
            return [sd * old_to_new[d // sd] + d % sd for d in v if d // sd in old_to_new]
        for lvl in self.levels:
            for nd in lvl:
                if nd.is_leaf:
                    nd.original_box = remap(nd.original_box)
                nd.active_box = remap(nd.active_box)
                nd.skel = remap(nd.skel)
                nd.redundant = remap(nd.redundant)
                nd.skelnear = remap(nd.skelnear)
        bursting = []
        for a in additions:
            leaf = self._leaf_of(float(new[a, 0]), float(new[a, 1]))
            leaf.original_box.extend(sd * a + j for j in range(sd))
            bursting.append(leaf)
            self.mark_neighbors_and_parents(leaf)
        for lvl in self.levels:
            for nd in lvl:
                nd.dofs_below = len(nd.original_box) if nd.is_leaf else 0
        for lv in range(len(self.levels) - 1, 0, -1):
            for nd in self.levels[lv]:
                nd.parent.dofs_below += nd.dofs_below
        sparse = []
        for d in deletions:
            x, y = float(old[d, 0]), float(old[d, 1])
            cur, marked = self.root, False
            while not cur.is_leaf:
                if cur.dofs_below < self.leaf_dofs and not marked:
                    marked = True
                    if not any(s is cur for s in sparse):
                        sparse.append(cur)
                midx, midy = cur.mid()
                cur = cur.children[_quadrant(x, y, midx, midy)]
            self.mark_neighbors_and_parents(cur)
        self.points = new.reshape(-1).copy()
        self._x = new[:, 0].tolist()
        self._y = new[:, 1].tolist()
        for nd in bursting:
            if nd.is_leaf and len(nd.original_box) > self.leaf_dofs:
                self._subdivide(nd)
        for nd in sparse:
            self.consolidate_node(nd)
        for nd in self.all_nodes():
            nd.neighbors = []
        self._neighbor_lists()
        self._number()

    # -- views used by tests ------------------------------------------------
    def all_nodes(self):
        for lvl in self.levels:
            yield from lvl

    def export(self):
        """Flat arrays in level order (the same layout the C-ABI exports)."""
        nodes = list(self.all_nodes())
        out = {
            "level_offsets": np.cumsum([0] + [len(l) for l in self.levels]).astype(np.int64),
            "parent": np.array([-1 if n.parent is None else n.parent.id for n in nodes], np.int32),
            "children": np.array([[(-1 if c is None else c.id) for c in n.children] for n in nodes],
                                 np.int32).reshape(-1, 4),
            "is_leaf": np.array([n.is_leaf for n in nodes], np.int32),
            "corners": np.array([n.corners for n in nodes], np.float64),
            "side": np.array([n.side for n in nodes], np.float64),
        }
        ob = [n.original_box for n in nodes]
        out["box_offsets"] = np.cumsum([0] + [len(b) for b in ob]).astype(np.int64)
        out["box_dofs"] = np.array([d for b in ob for d in b], np.int32)
        nb = [[m.id for m in n.neighbors] for n in nodes]
        out["nbr_offsets"] = np.cumsum([0] + [len(b) for b in nb]).astype(np.int64)
        out["nbrs"] = np.array([d for b in nb for d in b], np.int32)
        return out
