"""The per-box entry points of the C ABI against the oracle on the same inputs.

kigpu_kernel_block / kigpu_proxy_block (Kernel::operator(), get_proxy_mat,
kernel.cpp:99-150, :236-334), kigpu_interp_decomp (ki_Mat::id,
ki_mat.cpp:505-541) and kigpu_schur_eliminate (the arithmetic of
SkelFactorization::decouple, skel_factorization.cpp:55-91), driven with the
ID matrices and K_BN blocks of every compressed box of the oracle's
factorization.

Bars: kernel and proxy entries within 1e-13 of the block's largest entry
(same formulas, different rounding of the reciprocals); the ID's rank and
ordered skeleton identical, the redundant columns equal as a set and T
within 1e-8 relative column by column; X_rr within 1e-12 relative, dgetrf
pivots identical, L / U / the Schur update within 1e-9 relative."""
import numpy as np
import pytest

import kigpu
from oracle import geometry as G
from oracle import kernel as K
from oracle import linear_solve as LS
from oracle.kernel import RADIUS_RATIO
from oracle.skel import interpolative_decomposition

from helpers import star_problem


def rel(a, b):
    d = np.linalg.norm(np.asarray(a) - np.asarray(b))
    return d / max(np.linalg.norm(b), 1e-300)


def cases():
    b1 = star_problem(4096)
    # a star, not a circle: on the circle's symmetric boxes column norms tie
    # and the pivot order follows their rounding
    b2 = G.star(2048)
    b2.boundary_values = G.bc_tangent(b2)
    return [("laplace_star4096", b1, K.LAPLACE, 64), ("stokes_star2048", b2, K.STOKES, 64),
            ("neumann_star4096", b1, K.LAPLACE_NEUMANN, 64)]


@pytest.fixture(scope="module", params=cases(), ids=lambda c: c[0])
def factored(request):
    name, b, pde, leaf = request.param
    otree, okern, ofac = LS.factor(b, pde, 1e-6, leaf)
    boxes = [nd for nd in otree.all_nodes() if nd.compressed]
    assert boxes
    return name, b, pde, otree, okern, ofac, boxes


@pytest.mark.gpu
def test_kernel_and_proxy_blocks(ctx, factored):
    name, b, pde, otree, okern, ofac, boxes = factored
    print(f"[blocks] kernel/proxy rows {name}", flush=True)
    g = ctx.geometry(b.points, b.normals, b.weights, b.curvatures)
    nprox = 0
    for nd in boxes[::3]:
        box = nd.active_box
        rows, with_proxy = okern.id_rows(otree, nd)
        if rows:
            ref = okern(rows, box)
            got = g.block(pde, rows, box)
            assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
        if with_proxy:
            cx = nd.corners[0] + nd.side / 2.0
            cy = nd.corners[1] + nd.side / 2.0
            r = nd.side * RADIUS_RATIO
            ref = okern.proxy_block(cx, cy, r, box)
            got = g.proxy_block(pde, cx, cy, r, box)
            assert got.shape == ref.shape
            assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
            nprox += 1
    assert nprox > 0
    g.close()


@pytest.mark.gpu
def test_interp_decomp(ctx, factored):
    name, b, pde, otree, okern, ofac, boxes = factored
    print(f"[blocks] interp_decomp {name}: {len(boxes)} boxes", flush=True)
    for nd in boxes:
        A = okern.id_matrix(otree, nd)
        k, perm, T = interpolative_decomposition(A, 1e-6)
        assert k == len(nd.skel)          # the inputs are the ones the factor used
        kg, pg, Tg = ctx.interp_decomp(A, 1e-6)
        assert kg == k
        assert list(pg[:k]) == list(perm[:k])
        assert sorted(pg[k:]) == sorted(perm[k:])
        col = {c: j for j, c in enumerate(perm[k:])}
        for j, c in enumerate(pg[k:]):
            ref = T[:, col[c]]
            assert np.linalg.norm(Tg[:, j] - ref) <= 1e-8 * max(np.linalg.norm(ref), 1.0)


@pytest.mark.gpu
def test_schur_eliminate(ctx, factored):
    name, b, pde, otree, okern, ofac, boxes = factored
    print(f"[blocks] schur_eliminate {name}: {len(boxes)} boxes", flush=True)
    for nd in boxes:
        BN = nd.active_box
        KBN = okern(BN, BN) - ofac.schur_updates(BN, nd)
        k = nd.T.shape[0]
        res = ctx.schur_eliminate(KBN, k, nd.permutation, nd.T)
        assert res["status"] == 0
        s, r = np.asarray(nd.permutation[:k]), np.asarray(nd.permutation[k:])
        T = nd.T
        Xref = KBN[np.ix_(r, r)] - T.T @ KBN[np.ix_(s, r)] - (KBN[np.ix_(r, s)] - T.T @ KBN[np.ix_(s, s)]) @ T
        assert rel(res["X_rr"], Xref) <= 1e-12
        assert np.array_equal(res["X_rr_piv"], nd.X_piv + 1)   # dgetrf's, 1-based
        assert rel(res["X_rr_lu"], nd.X_lu) <= 1e-10
        assert rel(res["L"], nd.L) <= 1e-9
        assert rel(res["U"], nd.U) <= 1e-9
        assert rel(res["schur_update"], nd.schur) <= 1e-9


@pytest.mark.gpu
def test_entry_point_errors(ctx):
    print("[blocks] error paths", flush=True)
    rng = np.random.default_rng(4)
    with pytest.raises(kigpu.KigpuError):
        ctx.interp_decomp(rng.standard_normal((5, 9)), 1e-6)    # fewer rows than columns
    # a full-rank matrix is not compressible
    k, perm, T = ctx.interp_decomp(rng.standard_normal((40, 20)), 1e-6)
    assert k == 0 and perm is None
    # a zero redundant block: X_rr = 0 is singular
    n, k = 20, 12
    res = ctx.schur_eliminate(np.zeros((n, n)), k, np.arange(n), np.zeros((k, n - k)))
    assert res["status"] == 9
