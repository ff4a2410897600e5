"""In-domain masks and the domain evaluation on the GPU against the oracle.

kigpu_domain_mask runs Boundary::is_in_domain of every geometry family
(circle.cpp:34-42, annulus.cpp:66-78, donut.cpp:67-79,
cubic_boundary.cpp:182-230, the Fourier star) for all targets at once;
kigpu_domain_solution is the domain half of boundary_integral_solve
(linear_solve.cpp:143-180): forward apply, hole terms U_forward alpha and
the mask.

Bars: masks identical to the oracle's rules (the same comparisons in the
same order, with FMA contraction off) on random targets and on the 20 x 20
grid of the reference's tests; domain values within 1e-12 relative of the
oracle's dense forward (+ U_forward alpha) on a subset, exactly zero outside;
at 10^5 targets the values agree with the analytic solution to the solve's
accuracy (1e-6) and with the unmasked forward apply to 1e-14."""
import time

import numpy as np
import pytest

import kigpu
from oracle import cubic as C
from oracle import geometry as G
from oracle import kernel as K
from oracle import linear_solve as LS

from helpers import STOKES_SMALL_BORDER_FAULT, exact_point_source, gpu_tree, star_problem


def box_targets(b, n, seed, pad=0.05):
    lo, hi = b.xy().min(0) - pad, b.xy().max(0) + pad
    t = lo + G.uniform01(seed, 2 * n).reshape(-1, 2) * (hi - lo)
    g = np.linspace(0, 1, 20)
    grid = np.stack(np.meshgrid(lo[0] + g * (hi - lo)[0], lo[1] + g * (hi - lo)[1]), -1)
    return np.vstack([t, grid.reshape(-1, 2)])


def region_and_rule(name, b):
    if name in ("starfish", "ex1", "ex2", "ex3"):
        reg = kigpu.winding_region(b.points, b.num_outer, [h.num_nodes for h in b.holes])
        return reg, lambda t: C.in_winding_domain(b, t)
    if name == "star":
        reg = kigpu.star_region([0.3], [5.0], [0.0])
        return reg, lambda t: C.in_star_domain(t, [0.3], [5.0], [0.0])
    c, R = ((0.5, 0.5), 0.25) if name == "circle" else ((0.5, 0.5), 1.0)
    holes = [(h.center[0], h.center[1], h.radius) for h in b.holes]
    return kigpu.disc_region(c, R, holes), lambda t: C.in_disc_domain(t, c, R, b.holes)


GEOMS = [("starfish", lambda: C.starfish(8192), 20000), ("ex1", lambda: C.ex1(4096), 20000),
         ("ex2", lambda: C.ex2(4096), 20000), ("ex3", lambda: C.ex3(4096), 20000),
         ("circle", lambda: G.circle(1024), 100000), ("annulus", lambda: G.annulus(1000), 100000),
         ("donut", lambda: G.donut(1000), 100000), ("star", lambda: G.star(4096), 100000)]


@pytest.mark.gpu
@pytest.mark.parametrize("name,make,n", GEOMS, ids=[g[0] for g in GEOMS])
def test_mask_matches_oracle(ctx, name, make, n):
    print(f"[domain] mask {name} {n} targets", flush=True)
    b = make()
    reg, rule = region_and_rule(name, b)
    t = box_targets(b, n, 11)
    got = ctx.domain_mask(reg, t)
    want = rule(t)
    assert got.shape == want.shape
    assert np.array_equal(got, want), np.flatnonzero(got != want)[:10]
    assert 0.05 < got.mean() < 0.95


@pytest.mark.gpu
def test_mask_edge_cases(ctx):
    print("[domain] mask edge cases", flush=True)
    b = C.starfish(1024)
    reg, rule = region_and_rule("starfish", b)
    # targets on nodes, eps away from nodes, on node rows (y equal to a node's y)
    xy = b.xy()
    t = np.vstack([xy[:50], xy[:50] + [0.0099, 0.0], xy[:50] + [0.0101, 0.0],
                   np.stack([np.full(50, 0.5), xy[:50, 1]], 1)])
    assert np.array_equal(ctx.domain_mask(reg, t), rule(t))
    assert ctx.domain_mask(reg, np.zeros((0, 2))).shape == (0,)
    bad = kigpu.winding_region(b.points, 0)
    with pytest.raises(kigpu.KigpuError):
        ctx.domain_mask(bad, t)


@pytest.mark.gpu
def test_domain_solution_simply_connected_1e5_targets(ctx):
    # star N = 16384, leaf 64, tol 1e-10: values at 10^5 grid targets
    print("[domain] star 16384: 1e5 targets", flush=True)
    b = star_problem(16384)
    t = gpu_tree(ctx, b, 1, 64)
    t.factor(K.LAPLACE, 1e-10, kigpu.FP64)
    mu = t.solve(b.boundary_values)
    reg = kigpu.star_region([0.3], [5.0], [0.0])
    g = np.linspace(-1.3, 1.3, 317)
    tgt = np.stack(np.meshgrid(g, g), -1).reshape(-1, 2)
    assert tgt.shape[0] > 100000
    t0 = time.perf_counter()
    u = t.domain_solution(tgt, mu, region=reg)
    dt = time.perf_counter() - t0
    print(f"  {tgt.shape[0]} targets in {dt * 1e3:.1f} ms (incl. transfers)", flush=True)
    inside = C.in_star_domain(tgt, [0.3], [5.0], [0.0])
    assert np.all(u[~inside] == 0.0)
    ex = exact_point_source(tgt[inside])
    assert np.linalg.norm(u[inside] - ex) / np.linalg.norm(ex) <= 1e-6
    # the mask only zeroes: inside values are the forward apply's (a target
    # set of another size tiles the sources differently: last-bit changes)
    sub = np.flatnonzero(inside)[::50]
    fw = t.forward(tgt[sub], mu)
    assert np.linalg.norm(u[sub] - fw) <= 1e-14 * np.linalg.norm(fw)
    # and the forward apply is the oracle's dense forward
    kd = K.Kernel(1, K.LAPLACE, b, tgt[sub[:1500]].reshape(-1))
    ref = kd.forward() @ mu
    assert np.linalg.norm(u[sub[:1500]] - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.gpu
@pytest.mark.parametrize("pde", [
    pytest.param(K.LAPLACE, id="laplace"),
    pytest.param(K.STOKES, id="stokes", marks=pytest.mark.skip(reason=STOKES_SMALL_BORDER_FAULT))])
def test_domain_solution_with_holes(ctx, pde):
    # annulus (two holes): forward + U_forward alpha, masked, against the oracle
    print(f"[domain] annulus holes pde {pde}", flush=True)
    b = G.annulus(2048)
    sd = 2 if pde == K.STOKES else 1
    b.boundary_values = G.bc_single_electron(b) if sd == 1 else G.bc_tangent(b)
    t = gpu_tree(ctx, b, sd, 32)
    U = LS.initialize_U(pde, b.holes, b.points)
    Psi = LS.initialize_Psi(pde, b.holes, b)
    t.set_border(U, Psi)
    t.factor(pde, 1e-10, kigpu.FP64)
    mu = t.solve(b.boundary_values)
    alpha = t.last_alpha
    tgt = box_targets(b, 4000, 5)
    holes = [(h.center[0], h.center[1], h.radius) for h in b.holes]
    reg = kigpu.disc_region((0.5, 0.5), 1.0, holes)
    hc = [(h.center[0], h.center[1]) for h in b.holes]
    u = t.domain_solution(tgt, mu, alpha, hc, reg)
    inside = C.in_disc_domain(tgt, (0.5, 0.5), 1.0, b.holes)
    kd = K.Kernel(sd, pde, b, tgt.reshape(-1))
    ref = kd.forward() @ mu + LS.initialize_U(pde, b.holes, tgt.reshape(-1)) @ alpha
    ref = ref.reshape(-1, sd)
    ref[~inside] = 0.0
    got = u.reshape(-1, sd)
    assert np.all(got[~inside] == 0.0)
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
