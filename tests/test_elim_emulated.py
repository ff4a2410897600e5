"""The elimination of one box by one wavefront (decouple_wave_box, the body of
k_decouple_wave) and by one 1024-thread workgroup (decouple_block_box, the
body of k_decouple), both in kern-interp-old_amd/csrc/elim_dev.h, on the CPU,
on emulated GPU threads (tests/cpp/simt) under AddressSanitizer and UBSan:
every array it indexes -- geometry, pools, the factor slab, the children's
updates, the dynamic LDS -- is a heap block of exactly the box's size.

The box is decouple (skel_factorization.cpp:55-91) with the descendants'
updates accumulated child by child: with K = K(BN, BN) minus the children's
updates (block diagonal in active order), s the skeleton and r the
redundant DOFs, the slab must hold T, L = K_sn_r X_rr^{-1},
U = X_rr^{-1} K_r_sn, X_rr^{-1} and Acc = L K_r_sn plus the children's
updates on the skeleton, checked here against numpy on the same inputs:
Laplace and Stokes (two DOFs per point), leaves and boxes with children,
one to three register widths of the Gauss-Jordan, FP64 and FP32, ranks from
1 to n - 1, and an uncompressed box (its whole update handed up).
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import geometry as G
from oracle import kernel as K

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("elim") / "elim_emu")
    cmd = ["/opt/rocm/llvm/bin/clang++", "-std=c++20", "-O1", "-g", "-Wno-psabi",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer",
           "-I", os.path.join(ROOT, "tests", "cpp", "simt"),
           "-I", os.path.join(ROOT, "kern-interp-old_amd", "csrc"),
           "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "elim_emu.cpp"), "-o", exe, "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def make_box(pde, n, k, children, seed):
    """A box of n active DOFs of a star boundary: its DOFs (active order),
    the [skel | redundant] order as (DOF, active position) pairs, a random T
    and random symmetric children's updates of the given sizes."""
    rng = np.random.default_rng(seed)
    sd = 2 if pde == K.STOKES else 1
    b = G.star(512)
    npts = 512
    kern = K.Kernel(sd, pde, b)
    npb = (n + sd - 1) // sd
    span = max(64, npb + 16)
    start = int(rng.integers(0, npts - span))
    pts = start + np.sort(rng.choice(span, size=npb, replace=False))
    dofs = np.array([p * sd + c for p in pts for c in range(sd)])[:n]
    perm = rng.permutation(n)           # [skel | redundant] -> active position
    T = rng.standard_normal((k, n - k)) * 0.5 if k < n else np.zeros((0, 0))
    ups = []
    for sz in children:
        a = rng.standard_normal((sz, sz)) * 1e-3
        ups.append(a + a.T)
    assert not children or sum(children) == n
    return dict(b=b, kern=kern, sd=sd, npts=npts, dofs=dofs, perm=perm, T=T, ups=ups)


def run_box(emu, box, pde, n, k, children, tsize, rmax):
    b = box["b"]
    hdr = [tsize, rmax, pde, box["sd"], box["npts"], n, k, 0 if children else 1]
    hdr += list(children) + [0] * (4 - len(children))
    xy = b.points.reshape(-1, 2)
    nm = b.normals.reshape(-1, 2)
    geo = np.concatenate([xy[:, 0], xy[:, 1], nm[:, 0], nm[:, 1], b.weights, b.curvatures])
    sr = box["dofs"][box["perm"]].astype(np.int32)
    pos = box["perm"].astype(np.int32)
    inp = struct.pack("12i", *hdr) + geo.astype(np.float64).tobytes()
    inp += sr.tobytes() + pos.tobytes()
    inp += np.asfortranarray(box["T"]).astype(np.float64).tobytes(order="F")
    for a in box["ups"]:
        inp += np.asfortranarray(a).tobytes(order="F")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    r = subprocess.run([emu], input=inp, capture_output=True, env=env)
    assert r.returncode == 0, r.stderr.decode()[-4000:]
    status, osize = struct.unpack("2i", r.stdout[:8])
    slab = np.frombuffer(r.stdout[8:], np.float64)
    return status, osize, slab


def reference(box, n, k):
    dofs, perm, T = box["dofs"], box["perm"], box["T"]
    Kb = box["kern"](dofs, dofs)
    U = np.zeros((n, n))
    o = 0
    for a in box["ups"]:
        U[o:o + len(a), o:o + len(a)] = a
        o += len(a)
    Kb = Kb - U
    s, r = perm[:k], perm[k:]
    Kss, Ksr, Krs, Krr = (Kb[np.ix_(s, s)], Kb[np.ix_(s, r)], Kb[np.ix_(r, s)],
                          Kb[np.ix_(r, r)])
    K_r_sn = Krs - T.T @ Kss
    K_sn_r = Ksr - Kss @ T
    X = Krr - T.T @ Ksr - K_r_sn @ T
    Xi = np.linalg.inv(X)
    L = K_sn_r @ Xi
    Um = Xi @ K_r_sn
    Acc = L @ K_r_sn + U[np.ix_(s, s)]
    return L, Um, Xi, Acc


CASES = [
    # pde, n, k, children, tsize, rmax
    (K.LAPLACE, 20, 6, (), 8, 32),
    (K.LAPLACE, 30, 9, (12, 18), 8, 32),
    (K.LAPLACE, 64, 9, (16, 16, 16, 16), 8, 64),
    (K.LAPLACE, 63, 4, (31, 32), 4, 64),
    (K.LAPLACE, 17, 16, (9, 8), 4, 32),
    (K.STOKES, 64, 24, (20, 22, 22), 8, 64),
    (K.STOKES, 48, 30, (), 8, 64),
    (K.STOKES, 64, 50, (32, 32), 8, 64),
    (K.STOKES, 32, 10, (16, 16), 4, 32),
    (K.STOKES, 62, 61, (30, 32), 8, 64),
]


BLOCK_CASES = [
    # the workgroup kernel (rmax 0): boxes beyond the wavefront kernel's 64 DOFs too
    (K.LAPLACE, 20, 6, (), 8, 0),
    (K.LAPLACE, 64, 9, (16, 16, 16, 16), 4, 0),
    (K.LAPLACE, 130, 41, (30, 40, 0, 60), 8, 0),
    (K.STOKES, 96, 40, (), 8, 0),
    (K.STOKES, 150, 76, (40, 36, 38, 36), 4, 0),
    (K.STOKES, 62, 61, (30, 32), 8, 0),
]


@pytest.mark.parametrize("pde,n,k,children,tsize,rmax", CASES + BLOCK_CASES)
def test_elimination(emu, pde, n, k, children, tsize, rmax):
    box = make_box(pde, n, k, children, seed=n * 100 + k)
    status, osize, slab = run_box(emu, box, pde, n, k, children, tsize, rmax)
    assert status == 0 and osize == k
    r = n - k
    L, Um, Xi, Acc = reference(box, n, k)
    o = 0
    got = {}
    for name, shape in (("T", (k, r)), ("L", (k, r)), ("U", (r, k)), ("Xi", (r, r)),
                        ("Acc", (k, k))):
        sz = shape[0] * shape[1]
        got[name] = slab[o:o + sz].reshape(shape, order="F")
        o += sz
    tol = 1e-11 if tsize == 8 else 2e-4
    for name, ref in (("L", L), ("U", Um), ("Xi", Xi), ("Acc", Acc)):
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.abs(got[name] - ref).max() <= tol * scale * max(1.0, np.linalg.cond(Xi)), name


@pytest.mark.parametrize("rmax", [64, 0])
@pytest.mark.parametrize("pde,children", [(K.LAPLACE, (10, 14)), (K.STOKES, (20, 24))])
def test_uncompressed_box_hands_update_up(emu, pde, children, rmax):
    n = sum(children)
    box = make_box(pde, n, n, children, seed=7)
    status, osize, slab = run_box(emu, box, pde, n, n, children, 8, rmax)
    assert status == 0 and osize == n
    U = np.zeros((n, n))
    o = 0
    for a in box["ups"]:
        U[o:o + len(a), o:o + len(a)] = a
        o += len(a)
    assert np.array_equal(slab[:n * n].reshape(n, n, order="F"), U)
