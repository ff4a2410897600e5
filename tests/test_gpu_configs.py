"""The benchmarked workloads at their full sizes on the GPU.

* Laplace star r = 1 + 0.3 cos 5t, N = 65 536, leaf 64, id_tol 1e-6.
  FP64: per-box ranks, ordered skeleton lists and allskel equal to the
  oracle's; mu within 1e-9 of the oracle's; backward error <= 10 id_tol
  (kern_interp_test.cpp:90-93) from the GPU's FP64 direct sum; interior
  error <= 1e-6 against the point-source solution at 1 000 seeded targets
  (r < 0.6).  FP32 + refinement with 1, 16 and 128 right-hand sides (point
  sources at radius 2-4): the direct-sum residual of every column <= 1e-10,
  measured after the solve with kigpu_kernel_apply (the refinement itself
  uses the treecode residual), and the interior error of every column.
* Stokes double layer in the unit disc with 8 seeded circular holes, 16 384
  nodes on the outer circle and 2 048 per hole (65 536 DOFs, 24 border
  unknowns), Stokeslet data from (2, 1): FP64 ranks equal to the oracle's and
  [mu; alpha] within 1e-7 of the oracle's multiply_connected_solve (the bar
  of test_gpu_border.py); backward error of the bordered system
  [K U; Psi -I] from the GPU direct sum <= 10 id_tol; FP32 + refinement to
  1e-10; interior velocity against the Stokeslet (oracle FP64: 3.2e-7)."""
import numpy as np
import pytest

import kigpu
from oracle import geometry as G
from oracle import kernel as K
from oracle import linear_solve as LS

from helpers import (exact_point_source, gpu_tree, holed_disc_targets, oracle_ranks,
                     stokeslet_velocity, star_problem)

N = 65536


@pytest.fixture(scope="module")
def star_case():
    b = star_problem(N)
    otree, _, ofac = LS.factor(b, K.LAPLACE, 1e-6, 64)
    return b, otree, ofac


@pytest.mark.gpu
def test_star65536_fp64_parity(ctx, star_case):
    print("[configs] star N=65536 FP64 parity", flush=True)
    b, otree, ofac = star_case
    t = gpu_tree(ctx, b, 1, 64)
    t.factor(K.LAPLACE, 1e-6, kigpu.FP64)
    rank, active = t.ranks()
    assert np.array_equal(rank, oracle_ranks(otree))
    so, s, _, _ = t.lists()
    for i, nd in enumerate(otree.all_nodes()):
        if nd.compressed:
            assert active[i] == len(nd.active_box)
            assert list(s[so[i]:so[i + 1]]) == nd.skel
    assert list(t.allskel()) == otree.allskel
    f = b.boundary_values
    mu = t.solve(f)
    mu_ref = ofac.solve(otree, f)
    assert np.linalg.norm(mu - mu_ref) / np.linalg.norm(mu_ref) < 1e-9
    bwd = np.linalg.norm(t.apply(mu) - f) / np.linalg.norm(f)
    print(f"  backward error {bwd:.3e}", flush=True)
    assert bwd <= 10 * 1e-6
    tgt = G.interior_disc_targets(1000, 1234)
    ex = exact_point_source(tgt)
    err = np.linalg.norm(t.forward(tgt, mu) - ex) / np.linalg.norm(ex)
    print(f"  interior error {err:.3e}", flush=True)
    assert err <= 1e-6
    t.close()


@pytest.mark.gpu
@pytest.mark.parametrize("nrhs", [1, 16, 128])
def test_star65536_fp32_refine_rhs(ctx, star_case, nrhs):
    print(f"[configs] star N=65536 FP32+refine nrhs={nrhs}", flush=True)
    b = star_case[0]
    src = G.point_sources(nrhs, 77) if nrhs > 1 else np.array([[3.0, 3.0]])
    F = np.stack([exact_point_source(b.points, s) for s in src], 1)
    t = gpu_tree(ctx, b, 1, 64)
    t.factor(K.LAPLACE, 1e-6, kigpu.FP32_REFINE)
    mu = t.solve(F, refine_iters=20, rel_residual=1e-10).reshape(N, nrhs)
    assert t.last_info["rel_residual"] <= 1e-10
    R = t.apply(mu).reshape(N, nrhs) - F
    rel = np.linalg.norm(R, axis=0) / np.linalg.norm(F, axis=0)
    print(f"  direct-sum residual max {rel.max():.3e}, {t.last_info}", flush=True)
    assert rel.max() <= 1e-10
    tgt = G.interior_disc_targets(1000, 1234)
    U = t.forward(tgt, mu).reshape(-1, nrhs)
    for j, s in enumerate(src):
        ex = exact_point_source(tgt, s)
        assert np.linalg.norm(U[:, j] - ex) / np.linalg.norm(ex) <= 1e-6
    t.close()


@pytest.fixture(scope="module")
def stokes_case():
    b = G.multi_hole_disc(16384, 2048, 8, 7)
    b.boundary_values = G.bc_stokeslet(b)
    assert 2 * b.num_points == N
    otree, _, ofac = LS.factor(b, K.STOKES, 1e-6, 64)
    mu_ref, alpha_ref = ofac.multiply_connected_solve(otree, b.boundary_values)
    U = LS.initialize_U(K.STOKES, b.holes, b.points)
    Psi = LS.initialize_Psi(K.STOKES, b.holes, b)
    return b, otree, mu_ref, alpha_ref, U, Psi


def bordered_backward_error(t, b, U, Psi, mu, alpha):
    f = b.boundary_values
    r1 = t.apply(mu, K.STOKES) + U @ alpha - f
    r2 = Psi @ mu - alpha
    return np.sqrt(np.sum(r1 ** 2) + np.sum(r2 ** 2)) / np.linalg.norm(f)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [kigpu.FP64, kigpu.FP32_REFINE])
def test_stokes_holes_65536(ctx, stokes_case, precision):
    print(f"[configs] stokes 8 holes 65536 DOF precision {precision}", flush=True)
    b, otree, mu_ref, alpha_ref, U, Psi = stokes_case
    t = gpu_tree(ctx, b, 2, 64)
    t.set_border(U, Psi)
    t.factor(K.STOKES, 1e-6, precision)
    refine = 20 if precision == kigpu.FP32_REFINE else 0
    mu = t.solve(b.boundary_values, refine_iters=refine, rel_residual=1e-10)
    alpha = t.last_alpha
    bwd = bordered_backward_error(t, b, U, Psi, mu, alpha)
    tgt = holed_disc_targets(b, 500, 99)
    ex = stokeslet_velocity(tgt)
    u = t.forward(tgt, mu, K.STOKES) + LS.initialize_U(K.STOKES, b.holes, tgt) @ alpha
    err = np.linalg.norm(u - ex) / np.linalg.norm(ex)
    print(f"  backward error {bwd:.3e}  interior error {err:.3e}", flush=True)
    if precision == kigpu.FP64:
        rank, _ = t.ranks()
        assert np.array_equal(rank, oracle_ranks(otree))
        scale = np.linalg.norm(np.concatenate([mu_ref, alpha_ref]))
        assert np.linalg.norm(mu - mu_ref) / scale < 1e-7
        assert np.linalg.norm(alpha - alpha_ref) / scale < 1e-7
        assert bwd <= 10 * 1e-6
        assert err <= 1e-6
    else:
        assert t.last_info["rel_residual"] <= 1e-10
        assert bwd <= 1e-9
        assert err <= 1e-6
    t.close()
