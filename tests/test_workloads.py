"""The package's workload generators produce the same arrays as the oracle's
test geometries (the bench times the former, the tests check against the
latter)."""
import numpy as np

from kigpu import workloads as W
from oracle import geometry as G


def test_uniform_stream_matches():
    assert np.array_equal(W.uniform01(12345, 50), G.uniform01(12345, 50))


def test_star_and_random_star_match():
    for a, b in ((W.star(4096), G.star(4096)), (W.random_star(2048, 9), G.random_star(2048, 9))):
        for x, y in zip(a, (b.points, b.normals, b.weights, b.curvatures)):
            assert np.array_equal(x, y)


def test_bc_and_targets_match():
    b = G.star(1024)
    assert np.array_equal(W.point_source_bc(b.points), G.bc_point_source(b, (3.0, 3.0)))
    assert np.array_equal(W.interior_disc_targets(100, 1234), G.interior_disc_targets(100, 1234))
    assert np.array_equal(W.point_sources(7, 7), G.point_sources(7, 7))


def test_batch_sizes_range():
    n = W.batch_sizes(256, 4)
    assert n.min() >= 1024 and n.max() <= 16384 and np.all(n % 64 == 0)
    assert 6000 < n.mean() < 11000


def test_holed_disc_and_border_match():
    from oracle import kernel as K
    from oracle import linear_solve as LS
    (pts, nrm, w, kap), holes = W.holed_disc(4096, 512, 8, 7)
    b = G.multi_hole_disc(4096, 512, 8, 7)
    for x, y in zip((pts, nrm, w, kap), (b.points, b.normals, b.weights, b.curvatures)):
        assert np.array_equal(x, y)
    assert [(h[0], h[1], h[2]) for h in holes] == [(h.center[0], h.center[1], h.radius)
                                                  for h in b.holes]
    for stokes, pde in ((False, K.LAPLACE), (True, K.STOKES)):
        U, Psi = W.border_matrices(stokes, holes, pts, w)
        assert np.array_equal(U, LS.initialize_U(pde, b.holes, b.points))
        assert np.array_equal(Psi, LS.initialize_Psi(pde, b.holes, b))
    assert np.array_equal(W.stokeslet_bc(pts), G.bc_stokeslet(b))
