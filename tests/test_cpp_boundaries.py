"""The drop-in's boundary classes (host code of libkern_interp.so, no GPU)
against the oracle: CubicSpline / Ex1Boundary / Ex2Boundary / Ex3Boundary and
the analytic geometries node by node, their is_in_domain on the host, and
ki_Mat::real_eigenvalues against LAPACK's dgeev (numpy.linalg.eigvals).

Bars: nodes, normals and boundary data within 1e-13 (the spline solve is
the library's own LU instead of LAPACK's, so the knot derivatives may differ
in the last bit), weights and curvatures within 1e-12 relative; in-domain
answers identical; real eigenvalues within 1e-12 relative."""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import cubic as C
from oracle import geometry as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "kern-interp-old_amd")


@pytest.fixture(scope="module")
def dump(tmp_path_factory):
    assert os.path.exists(os.path.join(PKG, "libkern_interp.so")), "build first"
    exe = str(tmp_path_factory.mktemp("bd") / "boundary_dump")
    cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "boundary_dump.cpp"), "-o", exe, "-L" + PKG,
           "-lkern_interp", "-lkigpu", "-Wl,-rpath," + PKG]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def run(exe, args, stdin=b""):
    r = subprocess.run([exe, *map(str, args)], input=stdin, capture_output=True)
    assert r.returncode == 0, r.stderr.decode()[-2000:]
    return r.stdout


def geom(exe, name, N, bc, params=()):
    out = run(exe, ["geom", name, N, bc, *params])
    np_, nout, nh, nb = struct.unpack("4i", out[:16])
    o = 16
    holes = []
    for _ in range(nh):
        cx, cy, r = struct.unpack("3d", out[o:o + 24])
        (n,) = struct.unpack("i", out[o + 24:o + 28])
        holes.append(((cx, cy), r, n))
        o += 28
    arr = np.frombuffer(out[o:], np.float64)
    parts, k = [], 0
    for m in (2 * np_, 2 * np_, np_, np_, nb):
        parts.append(arr[k:k + m])
        k += m
    return nout, holes, parts


def oracle_geom(name, N, bc, params):
    if name in ("starfish", "ex1", "ex2", "ex3"):
        fn = getattr(C, name)
        return fn(N, bc) if name == "starfish" else fn(N, bc, list(params) or None)
    b = {"circle": G.circle, "annulus": G.annulus, "donut": G.donut}[name](N)
    b.num_outer = N
    f = np.zeros(C.bc_size(b, bc))
    C.apply_bc(b, f, 0, b.num_points, bc)
    b.boundary_values = f
    return b


CASES = [("starfish", 1024, "SINGLE_ELECTRON", ()), ("starfish", 8192, "TANGENT_VEC", ()),
         ("starfish", 1024, "ALL_ONES", ()), ("ex1", 4096, "DEFAULT", ()),
         ("ex1", 4096, "DEFAULT", (1.3,)), ("ex2", 4096, "DEFAULT", ()),
         ("ex2", 2048, "DEFAULT", (0.4, 0.6, 0.45, 0.55, 0.5, 0.35, 0.65, 0.5)),
         ("ex3", 4096, "EX3A", ()), ("ex3", 4096, "EX3B", (0.7, 0.7 + np.pi)),
         ("circle", 1024, "TANGENT_VEC", ()), ("annulus", 1000, "SINGLE_ELECTRON", ()),
         ("donut", 1000, "NORMAL_VEC", ())]


@pytest.mark.parametrize("name,N,bc,params", CASES)
def test_geometry_matches_oracle(dump, name, N, bc, params):
    nout, holes, (p, n, w, k, f) = geom(dump, name, N, bc, params)
    b = oracle_geom(name, N, bc, params)
    if name in ("starfish", "ex1", "ex2", "ex3"):
        assert nout == b.num_outer
    assert len(holes) == len(b.holes)
    for (c, r, m), h in zip(holes, b.holes):
        assert np.allclose(c, h.center, atol=1e-15) and r == h.radius and m == h.num_nodes
    assert p.shape == b.points.shape
    assert np.abs(p - b.points).max() < 1e-13
    assert np.abs(n - b.normals).max() < 1e-13
    assert np.abs(w - b.weights).max() < 1e-12 * np.abs(b.weights).max()
    assert np.abs(k - b.curvatures).max() < 1e-12 * np.abs(b.curvatures).max()
    assert np.abs(f - b.boundary_values).max() < 1e-13


@pytest.mark.parametrize("name,N,bc", [("starfish", 1024, "SINGLE_ELECTRON"),
                                       ("ex1", 4096, "DEFAULT"), ("ex3", 2048, "EX3A"),
                                       ("circle", 1024, "SINGLE_ELECTRON"),
                                       ("annulus", 1000, "SINGLE_ELECTRON"),
                                       ("donut", 1000, "SINGLE_ELECTRON")])
def test_is_in_domain_matches_oracle(dump, name, N, bc):
    b = oracle_geom(name, N, bc, ())
    lo, hi = b.xy().min(0) - 0.05, b.xy().max(0) + 0.05
    t = lo + G.uniform01(3, 6000).reshape(-1, 2) * (hi - lo)
    # and a 20 x 20 grid over the box as the reference's tests use
    g = np.linspace(0, 1, 20)
    grid = np.stack(np.meshgrid(lo[0] + g * (hi - lo)[0], lo[1] + g * (hi - lo)[1]), -1)
    t = np.vstack([t, grid.reshape(-1, 2)])
    out = run(dump, ["domain", name, N, bc],
              struct.pack("i", t.shape[0]) + np.ascontiguousarray(t).tobytes())
    got = np.frombuffer(out, np.uint8).astype(bool)
    if name in ("starfish", "ex1", "ex3"):
        want = C.in_winding_domain(b, t)
    else:
        c, R = ((0.5, 0.5), 0.25) if name == "circle" else ((0.5, 0.5), 1.0)
        want = C.in_disc_domain(t, c, R, b.holes)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n,seed", [(3, 1), (3, 2), (5, 3), (8, 4), (12, 5), (20, 6)])
def test_real_eigenvalues(dump, n, seed):
    A = np.random.default_rng(seed).standard_normal((n, n))
    out = run(dump, ["eig"], struct.pack("i", n) + np.asfortranarray(A).tobytes(order="F"))
    (m,) = struct.unpack("i", out[:4])
    got = np.sort(np.frombuffer(out[4:], np.float64))
    ev = np.linalg.eigvals(A)
    want = np.sort(ev[np.abs(ev.imag) < 1e-14].real)
    assert m == want.size
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(ev).max())


def test_cubic_roots(dump):
    # the companion matrix of (t - 1)(t + 2)(t - 0.5): roots 1, -2, 0.5
    coeffs = np.polynomial.polynomial.polyfromroots([1.0, -2.0, 0.5])   # c0 + c1 t + ...
    comp = np.zeros((3, 3))
    comp[1, 0] = comp[2, 1] = 1.0
    comp[:, 2] = -coeffs[:3] / coeffs[3]
    out = run(dump, ["eig"], struct.pack("i", 3) + np.asfortranarray(comp).tobytes(order="F"))
    got = np.sort(np.frombuffer(out[4:], np.float64))
    assert np.allclose(got, [-2.0, 0.5, 1.0], atol=1e-13)
