"""Solve sweeps replayed from captured HIP graphs (kigpu_set_graphs).

A tree captures its sweeps on the second solve with the same vector and
shape and replays them afterwards.  Every replay is checked bit for bit
against the direct launches of a context without graphs, with other device
work (another tree's factorization, applies) between the replays; a new
factorization of the tree drops its graphs and the replays follow it."""
import time

import numpy as np
import pytest

import kigpu
from oracle import geometry as G
from oracle import kernel as K
from oracle import linear_solve as LS

from helpers import gpu_tree, star_problem


@pytest.fixture(scope="module")
def gctx():
    c = kigpu.Context(0)
    c.set_graphs(True)
    yield c
    c.close()


def solves(ctx, b, pde, sd, leaf, tol, precision, rhs_list, border=None, refine=0, trans=False):
    t = gpu_tree(ctx, b, sd, leaf)
    if border is not None:
        t.set_border(*border)
    t.factor(pde, tol, precision)
    out = []
    for f in rhs_list:
        if trans:
            out.append(t.solve_transpose(f, refine_iters=refine))
        else:
            out.append(t.solve(f, refine_iters=refine))
    return t, out


@pytest.mark.gpu
@pytest.mark.parametrize("precision,nrhs,trans", [(kigpu.FP64, 1, False), (kigpu.FP64, 16, False),
                                                  (kigpu.FP32_REFINE, 1, False),
                                                  (kigpu.FP64, 1, True)])
def test_replays_match_direct_launches(ctx, gctx, precision, nrhs, trans):
    print(f"[graphs] star 16384 precision {precision} nrhs {nrhs} trans {trans}", flush=True)
    b = star_problem(16384)
    rng = np.random.default_rng(7)
    F = [rng.standard_normal((16384, nrhs)) if nrhs > 1 else rng.standard_normal(16384)
         for _ in range(5)]
    refine = 3 if precision == kigpu.FP32_REFINE else 0
    _, want = solves(ctx, b, K.LAPLACE, 1, 64, 1e-8, precision, F, refine=refine, trans=trans)
    tg = gpu_tree(gctx, b, 1, 64)
    tg.factor(K.LAPLACE, 1e-8, precision)
    other = gpu_tree(gctx, G.circle(2048), 1, 32)
    replays = 0
    for i, f in enumerate(F):
        x = tg.solve_transpose(f, refine_iters=refine) if trans else tg.solve(f, refine_iters=refine)
        assert np.array_equal(x, want[i]), f"solve {i}"
        replays += gctx.timings()["graph_replays"]
        # device work of another tree between the replays
        other.factor(K.LAPLACE, 1e-6, kigpu.FP64)
        other.apply(np.ones(2048))
    assert replays >= 3   # solves 2.. replay (the FP32 refinement replays its corrections too)
    # a new factorization drops the graphs: replays follow the new factors
    tg.factor(K.LAPLACE, 1e-6, precision)
    t2, want2 = solves(ctx, b, K.LAPLACE, 1, 64, 1e-6, precision, F[:3], refine=refine,
                       trans=trans)
    for i, f in enumerate(F[:3]):
        x = tg.solve_transpose(f, refine_iters=refine) if trans else tg.solve(f, refine_iters=refine)
        assert np.array_equal(x, want2[i]), f"solve {i} after refactor"


@pytest.mark.gpu
def test_bordered_replays(ctx, gctx):
    print("[graphs] annulus 2048 with border", flush=True)
    b = G.annulus(2048)
    b.boundary_values = G.bc_single_electron(b)
    border = (LS.initialize_U(K.LAPLACE, b.holes, b.points), LS.initialize_Psi(K.LAPLACE, b.holes, b))
    rng = np.random.default_rng(3)
    F = [rng.standard_normal(b.num_points) for _ in range(4)]
    t0, want = solves(ctx, b, K.LAPLACE, 1, 32, 1e-10, kigpu.FP64, F, border=border)
    alpha0 = []
    for f in F:
        t0.solve(f)
        alpha0.append(t0.last_alpha.copy())
    tg = gpu_tree(gctx, b, 1, 32)
    tg.set_border(*border)
    tg.factor(K.LAPLACE, 1e-10, kigpu.FP64)
    replays = 0
    for i, f in enumerate(F):
        assert np.array_equal(tg.solve(f), want[i])
        assert np.array_equal(tg.last_alpha, alpha0[i])
        replays += gctx.timings()["graph_replays"]
    assert replays >= 2


@pytest.mark.gpu
def test_bordered_replays_after_scratch_moves(ctx, gctx):
    # solves of 3, 3, 4, 3, 3 right-hand sides: the 4-column solve grows the
    # bordered step's scratch, so the 3-column graph captured before it is
    # not replayed (it would write the freed blocks): the next 3-column solve
    # runs directly and the one after captures again
    print("[graphs] annulus 2048 with border, 3/3/4/3/3 right-hand sides", flush=True)
    b = G.annulus(2048)
    b.boundary_values = G.bc_single_electron(b)
    border = (LS.initialize_U(K.LAPLACE, b.holes, b.points), LS.initialize_Psi(K.LAPLACE, b.holes, b))
    rng = np.random.default_rng(5)
    F = [rng.standard_normal((b.num_points, k)) for k in (3, 3, 4, 3, 3)]
    t0, want = solves(ctx, b, K.LAPLACE, 1, 32, 1e-10, kigpu.FP64, F, border=border)
    tg = gpu_tree(gctx, b, 1, 32)
    tg.set_border(*border)
    tg.factor(K.LAPLACE, 1e-10, kigpu.FP64)
    counts = []
    for i, f in enumerate(F):
        assert np.array_equal(tg.solve(f), want[i]), f"solve {i}"
        tm = gctx.timings()
        counts.append((tm["graph_captures"], tm["graph_replays"]))
    assert counts[1] == (1, 1) and counts[2] == (0, 0)
    assert counts[3] == (0, 0) and counts[4] == (1, 1)


@pytest.mark.gpu
def test_repeated_solve_time(ctx, gctx):
    # many solves with one factorization: direct launches against replays
    print("[graphs] star 65536: 50 solves, direct vs replayed", flush=True)
    b = star_problem(65536)
    f = b.boundary_values
    res = {}
    for name, c in (("direct", ctx), ("graph", gctx)):
        t = gpu_tree(c, b, 1, 64)
        t.factor(K.LAPLACE, 1e-6, kigpu.FP64)
        for _ in range(3):
            t.solve(f)
        t0 = time.perf_counter()
        dev = 0.0
        for _ in range(50):
            x = t.solve(f)
            dev += c.timings()["t_solve"]
        res[name] = (time.perf_counter() - t0) / 50
        res[name + "_dev"] = dev / 50
        res[name + "_x"] = x
    print(f"  per solve (host clock, incl. transfers): direct {res['direct'] * 1e3:.3f} ms, "
          f"graph {res['graph'] * 1e3:.3f} ms; on the device (copy in + sweeps): direct "
          f"{res['direct_dev'] * 1e3:.3f} ms, graph {res['graph_dev'] * 1e3:.3f} ms", flush=True)
    assert np.array_equal(res["direct_x"], res["graph_x"])
