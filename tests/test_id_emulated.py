"""The ID of one box as the ID kernels compute it (kern-interp-old_amd/csrc/
id_dev.h: id_box as k_id and k_id_wide run it, id_gram_box as k_id_gram runs
it) on the CPU, on emulated GPU threads (tests/cpp/simt) under
AddressSanitizer and UBSan, against the oracle's ID of the same box
(``Kernel.id_matrix`` + ``interpolative_decomposition``, ki_mat.cpp:505-541,
kernel.cpp:153-334).

Boxes come from the oracle's factorization of small stars, chosen to cover
every path: the register QR of k_id (<= 48 columns, <= 512 rows), its
trailing update in LDS and in the global workspace (the level-1/2 boxes of a
few thousand outside rows), both register shapes of k_id_wide (up to 64
columns and 512 rows; 48 columns and 640 rows, which the Stokes proxy block
of 512 rows reaches), and the Gram-matrix ID.  The Householder paths must
give the oracle's rank and ordered skeleton and T to rounding.  The Gram ID,
whose pivots FP64 resolves only to ~eps |A|^2, must give the same rank and
skeleton on all but near-tied boxes (none here); its T carries an error of
~eps cond(R11)^2, up to 1e-3 on these boxes, almost all of it along
directions that the skeleton columns map to nearly zero, so what is compared
is the interpolation A_skel T: within 1e-7 |A| of the oracle's (FP32 storage
of T is ~2e-8 of it), and the ID residual |A_red - A_skel T| within 1e-7 |A|
of the oracle's.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import geometry as G
from oracle import kernel as K
from oracle import linear_solve as LS
from oracle import skel as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LDS_CAP = 163840
PROXY = 128


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("id") / "id_emu")
    cmd = ["/opt/rocm/llvm/bin/clang++", "-std=c++20", "-O1", "-g", "-Wno-psabi",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer",
           "-I", os.path.join(ROOT, "tests", "cpp", "simt"),
           "-I", os.path.join(ROOT, "kern-interp-old_amd", "csrc"),
           "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "id_emu.cpp"), "-o", exe, "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def a16(x):
    return (x + 15) & ~15


def reg_box(m, n):
    return n <= 3 * 16 and m <= 8 * 64


def wide_box(m, n):
    return not reg_box(m, n) and ((n <= 64 and m <= 512) or (n <= 48 and m <= 640))


def id_lds(m, n, q, in_lds):
    """id_lds_layout (id_dev.h) in bytes, FP64 matrix."""
    o = a16(m * n * 8) if in_lds else 0
    for b in (48 * q, 48 * n, 4 * q, 4 * n, 8 * n, 8 * n, 4 * n):
        o = a16(o + b)
    if reg_box(m, n) or wide_box(m, n):
        o = a16(o + 8 * (m + 2))
    return a16(o + 8 * 32)


_CACHE = {}


def factored(name):
    """Oracle factorization of one of the test problems, with its boxes."""
    if name not in _CACHE:
        if name == "laplace":
            b, pde, tol, leaf = G.star(4096), K.LAPLACE, 1e-6, 64
        elif name == "neumann":
            b, pde, tol, leaf = G.star(2048), K.LAPLACE_NEUMANN, 1e-6, 64
        else:
            b, pde, tol, leaf = G.star(2048), K.STOKES, 1e-6, 64
        tree, kern, fac = LS.factor(b, pde, tol, leaf)
        boxes = []
        for nd in tree.all_nodes():
            if nd.level == 0 or len(nd.active_box) < 16:
                continue
            rows, proxy = kern.id_rows(tree, nd)
            m = 2 * len(rows) + ((4 if pde == K.STOKES else 2) * PROXY if proxy else 0)
            if m == 0:
                continue
            boxes.append((nd, rows, proxy, m))
        _CACHE[name] = (b, pde, tol, tree, kern, boxes)
    return _CACHE[name]


def run_box(emu, kind, tsize, name, nd, rows, proxy):
    b, pde, tol, tree, kern, _ = _CACHE[name]
    sd = 2 if pde == K.STOKES else 1
    n, q = len(nd.active_box), len(rows)
    m = 2 * q + ((4 if pde == K.STOKES else 2) * PROXY if proxy else 0)
    in_lds = int(id_lds(m, n, q, True) <= LDS_CAP) if kind != 2 else 1
    cx = nd.corners[0] + nd.side / 2.0
    cy = nd.corners[1] + nd.side / 2.0
    rad = nd.side * 1.5
    pw = 2.0 * np.pi * rad / PROXY
    xy = b.points.reshape(-1, 2)
    nm = b.normals.reshape(-1, 2)
    npts = xy.shape[0]
    inp = struct.pack("9i", kind, tsize, pde, sd, npts, n, q, int(proxy), in_lds)
    inp += struct.pack("5d", tol, cx, cy, rad, pw)
    inp += np.concatenate([xy[:, 0], xy[:, 1], nm[:, 0], nm[:, 1], b.weights,
                           b.curvatures]).astype(np.float64).tobytes()
    inp += np.asarray(nd.active_box, np.int32).tobytes() + np.asarray(rows, np.int32).tobytes()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    r = subprocess.run([emu], input=inp, capture_output=True, env=env)
    assert r.returncode == 0, r.stderr.decode()[-4000:]
    status, rank = struct.unpack("2i", r.stdout[:8])
    piv = np.frombuffer(r.stdout[8:8 + 4 * n], np.int32)
    T = np.frombuffer(r.stdout[8 + 4 * n:], np.float64)
    return status, rank, piv, T, in_lds


def pick(name, pred, count):
    _, _, _, _, _, boxes = factored(name)
    chosen = [bx for bx in boxes if pred(len(bx[0].active_box), bx[3], bx[0].level)]
    if len(chosen) > count:  # spread over the candidates
        idx = np.linspace(0, len(chosen) - 1, count).round().astype(int)
        chosen = [chosen[i] for i in idx]
    return chosen


def check(emu, kind, tsize, name, nd, rows, proxy, gram=False):
    b, pde, tol, tree, kern, _ = _CACHE[name]
    A = kern.id_matrix(tree, nd)
    k, perm, T = S.interpolative_decomposition(A, tol)
    status, rank, piv, Tg, _ = run_box(emu, kind, tsize, name, nd, rows, proxy)
    assert status == 0
    n = len(nd.active_box)
    if k == 0:
        assert rank == -1
        return
    assert rank == k, f"rank {rank} != {k} (n {n}, m {A.shape[0]})"
    assert list(piv[:k]) == list(perm[:k]), "skeleton"
    Tg = Tg.reshape(k, n - k, order="F")
    # T's columns follow the redundant order, which is the QR's up to the
    # rank step; compare by redundant column
    red = {c: j for j, c in enumerate(perm[k:])}
    Tr = np.empty_like(Tg)
    for j, c in enumerate(piv[k:]):
        Tr[:, j] = T[:, red[c]]
    if gram:
        As, Ar, na = A[:, list(perm[:k])], A[:, list(piv[k:])], np.linalg.norm(A)
        assert np.linalg.norm(As @ (Tg - Tr)) <= 1e-7 * na, "interpolation"
        assert np.linalg.norm(Ar - As @ Tg) <= np.linalg.norm(Ar - As @ Tr) + 1e-7 * na
        return
    scale = max(np.abs(Tr).max(), 1.0)
    bar = 1e-9 if tsize == 8 else 1e-5
    assert np.abs(Tg - Tr).max() <= bar * scale


@pytest.mark.parametrize("name", ["laplace", "stokes"])
def test_k_id_register_qr(emu, name):
    for nd, rows, proxy, m in pick(name, lambda n, m, l: reg_box(m, n), 3):
        check(emu, 0, 8, name, nd, rows, proxy)


def test_k_id_trailing_update_in_lds_and_workspace(emu):
    # boxes of more than 48 columns: in LDS when the matrix fits, else (the
    # level-1/2 boxes of outside rows) in the global workspace
    lds = pick("laplace", lambda n, m, l: not reg_box(m, n) and not wide_box(m, n)
               and id_lds(m, n, 0, True) <= LDS_CAP, 2)
    glb = pick("laplace", lambda n, m, l: not reg_box(m, n) and not wide_box(m, n)
               and id_lds(m, n, 0, True) > LDS_CAP, 2)
    assert lds and glb
    for nd, rows, proxy, m in lds + glb:
        check(emu, 0, 4, "laplace", nd, rows, proxy)


@pytest.mark.parametrize("name,shape", [("laplace", 8), ("stokes", 10)])
def test_k_id_wide(emu, name, shape):
    # shape 8: up to 64 columns and 512 rows; 10: 48 columns and 640 rows
    pred = (lambda n, m, l: wide_box(m, n) and m <= 512) if shape == 8 else \
        (lambda n, m, l: wide_box(m, n) and m > 512)
    boxes = pick(name, pred, 3)
    assert boxes
    for nd, rows, proxy, m in boxes:
        check(emu, 1, 8, name, nd, rows, proxy)


@pytest.mark.parametrize("name", ["laplace", "neumann", "stokes"])
def test_gram_id(emu, name):
    for nd, rows, proxy, m in pick(name, lambda n, m, l: n <= 64, 3):
        check(emu, 2, 4, name, nd, rows, proxy, gram=True)
