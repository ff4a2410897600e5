"""Incremental re-factorization edits on the host: ``QuadTree::perturb``,
``copy_into``, ``mark_neighbors_and_parents`` and ``consolidate_node``
(quadtree.cpp:288-641) of the drop-in library
(kern-interp-old_amd/src/quadtree.cpp) against the oracle's restatement
(``oracle.tree.QuadTree.perturb``), and both against a perturbation traced
by hand from the reference's code (tests/golden/perturb_trace_4pts.md).

The C++ side runs on the CPU (tests/cpp/perturb_host.cpp, built with ASan
and UBSan): the tree is built on the host by ``recursive_add`` and
``compute_neighbor_lists``, optionally marked as factored (every node
compressed, synthetic skeleton lists), copied with ``copy_into``, and the
copy perturbed; the device rebuild that ends ``perturb`` goes to a test
double of ``kigpu_tree_import`` that records the arrays it is given.

Compared: the level order of the nodes, parents, children, neighbor lists
in order, corners, leaf and compression flags, DOFs-below counts and every
DOF list.  The reference keeps empty trailing levels after a merge; the
drop-in drops them before the device rebuild, so the oracle's are stripped
for the comparison.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import geometry as G
from oracle.tree import QuadTree

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("perturb") / "perturb_host")
    src = os.path.join(ROOT, "kern-interp-old_amd", "src")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-I", os.path.join(ROOT, "include"), "-I", src,
           os.path.join(ROOT, "tests", "cpp", "perturb_host.cpp"),
           os.path.join(src, "quadtree.cpp"), os.path.join(src, "ki_mat.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def run_host(exe, sd, leaf, factored, old, new):
    old = np.asarray(old, np.float64).reshape(-1)
    new = np.asarray(new, np.float64).reshape(-1)
    txt = f"{sd} {leaf} {int(factored)}\n{old.size // 2}\n"
    txt += " ".join(repr(float(v)) for v in old) + f"\n{new.size // 2}\n"
    txt += " ".join(repr(float(v)) for v in new) + "\n"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1")
    r = subprocess.run([exe], input=txt, capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    return json.loads(r.stdout)


def mark_factored(t):
    """The harness's synthetic factor state, on the oracle's tree."""
    for nd in t.all_nodes():
        nd.compressed = True
        nd.active_box = list(nd.original_box)
        h = len(nd.active_box) // 2
        nd.skel = nd.active_box[:h]
        nd.redundant = nd.active_box[h:]
        nd.skelnear = list(nd.skel)


def oracle_view(t):
    """The oracle's tree in the harness's JSON layout (empty trailing levels
    stripped)."""
    levels = [list(l) for l in t.levels]
    while levels and not levels[-1]:
        levels.pop()
    pos = {}
    for l, lv in enumerate(levels):
        for i, nd in enumerate(lv):
            pos[id(nd)] = [l, i]

    def p(nd):
        return None if nd is None else pos[id(nd)]
    out = []
    for lv in levels:
        row = []
        for nd in lv:
            row.append({"level": nd.level, "leaf": int(nd.is_leaf),
                        "compressed": int(nd.compressed), "dofs_below": nd.dofs_below,
                        "parent": p(nd.parent), "children": [p(c) for c in nd.children],
                        "neighbors": [p(m) for m in nd.neighbors],
                        "corners": list(nd.corners), "side": nd.side,
                        "original_box": list(nd.original_box),
                        "active_box": list(nd.active_box), "skel": list(nd.skel),
                        "redundant": list(nd.redundant), "skelnear": list(nd.skelnear)})
        out.append(row)
    return out


def assert_same_tree(got, want):
    assert len(got) == len(want), "number of levels"
    for l, (gl, wl) in enumerate(zip(got, want)):
        assert len(gl) == len(wl), f"level {l}: node count"
        for i, (g, w) in enumerate(zip(gl, wl)):
            for k in w:
                assert g[k] == w[k], f"level {l} node {i}: {k}: {g[k]} != {w[k]}"


def check_against_fixture(view, fx, sd, leaf):
    names = fx["names"]
    lo = fx["level_offsets"]
    flat = [nd for lv in view for nd in lv]
    assert [len(lv) for lv in view] == list(np.diff(lo)), "level sizes"
    ids = {}
    for l, lv in enumerate(view):
        for i, _ in enumerate(lv):
            ids[(l, i)] = lo[l] + i
    for x, nd in enumerate(flat):
        name = names[x]
        par = -1 if nd["parent"] is None else ids[tuple(nd["parent"])]
        assert par == fx["parent"][x], f"{name}: parent"
        kids = [-1 if c is None else ids[tuple(c)] for c in nd["children"]]
        assert kids == fx["children"][x], f"{name}: children"
        assert nd["leaf"] == fx["is_leaf"][x], f"{name}: leaf"
        box = [d for q in fx["original_box_points"][x] for d in range(sd * q, sd * q + sd)]
        assert nd["original_box"] == box, f"{name}: original box"
        if fx["dofs_below_points"][x] is not None:
            assert nd["dofs_below"] == sd * fx["dofs_below_points"][x], f"{name}: dofs below"
        assert nd["compressed"] == fx["compressed"][x], f"{name}: compressed"
        if "lists_dofs" in fx:
            for key, field in (("active", "active_box"), ("skel", "skel"),
                               ("redundant", "redundant"), ("skelnear", "skelnear")):
                assert nd[field] == fx["lists_dofs"][str(sd)][x][key], f"{name}: {field}"


@pytest.fixture(scope="module")
def fixture():
    with open(os.path.join(GOLDEN, "perturb_trace_4pts.json")) as fh:
        return json.load(fh)


@pytest.mark.parametrize("edit", [0, 1])
def test_oracle_perturb_matches_hand_trace(fixture, edit):
    fx = fixture["perturbations"][edit]
    for case in fx["cases"]:
        t = QuadTree(fixture["old_points"], case["sd"], case["leaf_dofs"])
        mark_factored(t)
        t.perturb(fx["new_points"])
        check_against_fixture(oracle_view(t), fx, case["sd"], case["leaf_dofs"])


@pytest.mark.parametrize("edit", [0, 1])
def test_host_perturb_matches_hand_trace(harness, fixture, edit):
    fx = fixture["perturbations"][edit]
    for case in fx["cases"]:
        out = run_host(harness, case["sd"], case["leaf_dofs"], True, fixture["old_points"],
                       fx["new_points"])
        check_against_fixture(out["after"], fx, case["sd"], case["leaf_dofs"])


def moved_hole_case(n, m, shift, seed):
    """A circle of n nodes with a hole of m nodes; the hole moves by `shift`
    (its nodes are deleted and re-added elsewhere) and a few outer nodes
    are dropped: additions, deletions, bursting leaves and merges."""
    rng = np.random.default_rng(seed)
    th = 2 * np.pi * np.arange(n) / n
    outer = np.stack([0.5 + np.cos(th), 0.5 + np.sin(th)], 1)
    ph = 2 * np.pi * np.arange(m) / m
    hole = np.stack([0.6 + 0.15 * np.cos(ph), 0.4 + 0.15 * np.sin(ph)], 1)
    old = np.concatenate([outer, hole])
    keep = np.ones(n, bool)
    keep[rng.choice(n, size=n // 10, replace=False)] = False
    new = np.concatenate([outer[keep], hole + np.asarray(shift)])
    return old, new


CASES = [
    # sd, leaf, old/new generator
    (1, 8, lambda: moved_hole_case(256, 64, (-0.3, 0.25), 1)),
    (1, 32, lambda: moved_hole_case(1024, 256, (-0.35, 0.3), 2)),
    (2, 16, lambda: moved_hole_case(512, 128, (0.1, 0.35), 3)),
    (1, 4, lambda: (G.star(300).points.reshape(-1, 2),
                    np.concatenate([G.star(300).points.reshape(-1, 2)[::3],
                                    0.2 * G.star(90).points.reshape(-1, 2)]))),
    (2, 6, lambda: (G.circle(200).points.reshape(-1, 2),
                    G.circle(200).points.reshape(-1, 2)[:20])),   # most points deleted
]


@pytest.mark.parametrize("sd,leaf,gen", CASES)
def test_host_perturb_matches_oracle(harness, sd, leaf, gen):
    old, new = gen()
    out = run_host(harness, sd, leaf, True, old, new)
    t = QuadTree(old.reshape(-1), sd, leaf)
    before = oracle_view(t)
    mark_factored(t)
    assert_same_tree(out["before"], oracle_view(t))
    t.perturb(new.reshape(-1))
    assert_same_tree(out["after"], oracle_view(t))
    assert before  # the unfactored build agrees too (checked through "before")


@pytest.mark.parametrize("sd,leaf,gen", CASES[:3])
def test_copy_into_is_a_deep_copy(harness, sd, leaf, gen):
    old, new = gen()
    out = run_host(harness, sd, leaf, True, old, new)
    assert out["copy"] == out["before"]
    assert out["copy_shares_nodes"] == 0


@pytest.mark.parametrize("sd,leaf,gen", CASES)
def test_device_rebuild_gets_the_edited_tree(harness, sd, leaf, gen):
    # the arrays perturb hands to kigpu_tree_import: nodes in level order,
    # parents, children, leaves' boxes, neighbor lists
    old, new = gen()
    out = run_host(harness, sd, leaf, True, old, new)
    after = out["after"]
    imp = out["import"]
    assert imp["nlevels"] == len(after)
    assert imp["loff"] == list(np.cumsum([0] + [len(lv) for lv in after]))
    flat = [nd for lv in after for nd in lv]
    ids = {}
    for l, lv in enumerate(after):
        for i, _ in enumerate(lv):
            ids[(l, i)] = imp["loff"][l] + i
    for x, nd in enumerate(flat):
        assert imp["parent"][x] == (-1 if nd["parent"] is None else ids[tuple(nd["parent"])])
        kids = [-1 if c is None else ids[tuple(c)] for c in nd["children"]]
        assert imp["child"][4 * x:4 * x + 4] == (kids if not nd["leaf"] else [-1] * 4)
        box = imp["dofs"][imp["boff"][x]:imp["boff"][x + 1]]
        assert box == (nd["original_box"] if nd["leaf"] else [])
        nb = imp["nbrs"][imp["noff"][x]:imp["noff"][x + 1]]
        assert nb == [ids[tuple(p)] for p in nd["neighbors"]]
