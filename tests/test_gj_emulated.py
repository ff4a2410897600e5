"""The Gauss-Jordan inverses of kern-interp-old_amd/csrc run on the CPU on
emulated GPU threads (tests/cpp/simt/hip/hip_runtime.h), built with
AddressSanitizer and UBSan: every array the routines index (matrix, pivot
and permutation vectors, LDS scratch vectors, workspaces) is a heap block of
exactly its kernel size, so an index past any of them fails the run.

They replace dgetrf + dgetri / dgetrs of the reference's root, bordered and
X_rr LUs (ki_mat.cpp:427-493, skel_factorization.cpp:98-128, :236-242) with
the same pivot order, so the checks are:

* block_gj_inverse_perm (device_util.h, the workgroup elimination's X_rr
  inverse) against LAPACK's inverse (numpy.linalg.inv) on well-conditioned
  matrices, and its pivot sequence against dgetrf's;
* the register-resident Gauss-Jordan of the one-wavefront elimination
  (wave_gj_reg: row i in lane i, no row interchanges, the pivot row
  broadcast through LDS) in FP64 and FP32 at every register width it has
  (16, 32, 64 rows), against LAPACK's inverse;
* a column without a usable pivot (all zero, or NaN): each routine reports
  it and every index stays in bounds.  With the search of earlier code
  (best index starting at n) a zero pivot turned the next column into NaN,
  the search returned n, and row n, piv[c] = n, sigma[n] and column n were
  read and written past the arrays;
* the blocked Gauss-Jordan of dense_inv.h (the root and bordered
  inverses, k_dense_gj), with its matrix in LDS and in the global
  workspace, one wave or four waves on the panel, batched matrices of
  different orders, and the root matrix allskel_mat and bordered matrix S
  of the annulus at id_tol 1e-13 (kern_interp_test.cpp:413-430) from the
  oracle: allskel_mat is numerically singular there (two pivots at
  roundoff level, cond ~4e17), so the routine must stay inside its arrays,
  and S's inverse must be as accurate as cond(S) allows.  Built with
  clang++ (the emulated matrix-core step takes a clang vector).
"""
import os
import struct
import subprocess

import numpy as np
import pytest
from scipy.linalg import lapack

from oracle import geometry as G
from oracle import kernel as K
from oracle import linear_solve as LS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PERM, REG64, REG32 = 0, 3, 4


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("gj") / "gj_emu")
    cmd = ["g++", "-std=c++20", "-O2", "-g", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
           "-I", os.path.join(ROOT, "tests", "cpp", "simt"),
           "-I", os.path.join(ROOT, "kern-interp-old_amd", "csrc"),
           "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "gj_emu.cpp"), "-o", exe, "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def run(exe, A, variant, threads):
    n = A.shape[0]
    inp = struct.pack("3i", variant, n, threads) + np.asfortranarray(A).tobytes(order="F")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    r = subprocess.run([exe], input=inp, capture_output=True, env=env)
    assert r.returncode == 0, r.stderr.decode()[-4000:]
    out = r.stdout
    ok = struct.unpack("i", out[:4])[0]
    piv = np.frombuffer(out[4:4 + 4 * n], np.int32)
    inv = np.frombuffer(out[4 + 4 * n:], np.float64).reshape(n, n, order="F")
    return bool(ok), piv, inv


@pytest.mark.parametrize("variant,threads", [(PERM, 256), (PERM, 1024)])
def test_well_conditioned(emu, variant, threads):
    rng = np.random.default_rng(11)
    n = 70
    A = rng.standard_normal((n, n)) + 4 * np.eye(n)
    ok, piv, inv = run(emu, A, variant, threads)
    assert ok
    ref = np.linalg.inv(A)
    assert np.abs(inv - ref).max() <= 1e-12 * np.abs(ref).max()
    _, lpiv, info = lapack.dgetrf(A)
    assert info == 0
    assert np.array_equal(piv, lpiv)   # dgetrf's pivot order (0-based)


@pytest.fixture(scope="module")
def annulus_tight():
    b = G.annulus(4096)
    b.boundary_values = G.bc_single_electron(b)
    otree, _, _ = LS.factor(b, K.LAPLACE, 1e-13, 32)
    return otree


@pytest.mark.parametrize("variant,threads", [(PERM, 256), (PERM, 1024)])
@pytest.mark.parametrize("bad", ["zero", "nan"])
def test_unusable_pivot_reported_in_bounds(emu, variant, threads, bad):
    rng = np.random.default_rng(5)
    n = 50
    A = rng.standard_normal((n, n))
    A[:, 5] = 0.0 if bad == "zero" else np.nan
    ok, piv, _ = run(emu, A, variant, threads)
    assert not ok
    assert np.all((piv >= 0) & (piv < n))


@pytest.mark.parametrize("variant", [REG64, REG32])
@pytest.mark.parametrize("r", [1, 7, 16, 17, 25, 32, 33, 59, 64])
def test_register_gauss_jordan(emu, variant, r):
    rng = np.random.default_rng(100 + r)
    X = rng.standard_normal((r, r)) + 3 * np.eye(r)
    ok, piv, Z = run(emu, X, variant, 64)
    assert ok
    assert sorted(piv) == list(range(r))   # each row pivots once
    ref = np.linalg.inv(X)
    tol = 1e-12 if variant == REG64 else 2e-5
    assert np.abs(Z - ref).max() <= tol * np.abs(ref).max()


@pytest.mark.parametrize("variant", [REG64, REG32])
@pytest.mark.parametrize("bad", ["zero", "nan"])
def test_register_gauss_jordan_unusable_pivot(emu, variant, bad):
    rng = np.random.default_rng(9)
    r = 40
    X = rng.standard_normal((r, r))
    X[:, 7] = 0.0 if bad == "zero" else np.nan
    ok, piv, _ = run(emu, X, variant, 64)
    assert not ok
    assert np.all((piv >= 0) & (piv < 64))


# --- blocked Gauss-Jordan (dense_inv.h) -------------------------------------

@pytest.fixture(scope="module")
def dgj(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("dgj") / "dense_gj_emu")
    cmd = ["/opt/rocm/llvm/bin/clang++", "-std=c++20", "-O2", "-g", "-Wno-psabi",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer",
           "-I", os.path.join(ROOT, "tests", "cpp", "simt"),
           "-I", os.path.join(ROOT, "kern-interp-old_amd", "csrc"),
           "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "dense_gj_emu.cpp"), "-o", exe, "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def run_blocked(exe, mats, in_lds):
    ns = [m.shape[0] for m in mats]
    inp = struct.pack("2i", int(in_lds), len(mats)) + struct.pack(f"{len(ns)}i", *ns)
    inp += b"".join(np.asfortranarray(m).tobytes(order="F") for m in mats)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    r = subprocess.run([exe], input=inp, capture_output=True, env=env)
    assert r.returncode == 0, r.stderr.decode()[-4000:]
    out = r.stdout
    status = struct.unpack("i", out[:4])[0]
    sn = sum(ns)
    piv = np.frombuffer(out[4:4 + 4 * sn], np.int32)
    inv, off = [], 4 + 4 * sn
    for n in ns:
        inv.append(np.frombuffer(out[off:off + 8 * n * n], np.float64).reshape(n, n, order="F"))
        off += 8 * n * n
    return status, np.split(piv, np.cumsum(ns)[:-1]), inv


def dgetrf_rows(A):
    # the row each dgetrf step pivots on, as an index of the original matrix
    _, lpiv, info = lapack.dgetrf(A)
    perm = np.arange(A.shape[0])
    for c, p in enumerate(lpiv):
        perm[[c, p]] = perm[[p, c]]
    return perm


@pytest.mark.parametrize("n,in_lds", [(1, True), (5, True), (16, True), (17, False), (70, True),
                                      (70, False), (257, False)])
def test_blocked_gj_well_conditioned(dgj, n, in_lds):
    rng = np.random.default_rng(20 + n)
    A = rng.standard_normal((n, n)) + 4 * np.eye(n)
    status, piv, inv = run_blocked(dgj, [A], in_lds)
    assert status == 0
    ref = np.linalg.inv(A)
    assert np.abs(inv[0] - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.array_equal(piv[0], dgetrf_rows(A))   # dgetrf's pivots, as rows


def test_blocked_gj_batched_orders(dgj):
    # the root launch: instances of different orders back to back, one empty
    rng = np.random.default_rng(8)
    mats = [rng.standard_normal((n, n)) + 3 * np.eye(n) for n in (33, 0, 48, 7)]
    status, piv, inv = run_blocked(dgj, mats, True)
    assert status == 0
    for A, X, p in zip(mats, inv, piv):
        if A.shape[0]:
            assert np.abs(X @ A - np.eye(A.shape[0])).max() <= 1e-12
            assert np.array_equal(p, dgetrf_rows(A))


def test_blocked_gj_root_annulus_tol_1e13(dgj, annulus_tight):
    # the root of the bordered annulus at id_tol 1e-13 (n = 217): numerically
    # singular, so the check is that every index stays in bounds and the
    # result is finite, with the pivots dgetrf takes
    A = annulus_tight.allskel_mat
    status, piv, inv = run_blocked(dgj, [A], False)
    assert status == 0
    assert np.isfinite(inv[0]).all()
    assert sorted(piv[0]) == list(range(A.shape[0]))


def test_blocked_gj_bordered_annulus_tol_1e13(dgj, annulus_tight):
    S = annulus_tight.S_mat
    status, _, inv = run_blocked(dgj, [S], False)
    assert status == 0
    ref = np.linalg.inv(S)
    bound = 1e-15 * np.linalg.cond(S, 1) * 100
    assert np.abs(inv[0] - ref).max() <= bound * np.abs(ref).max()


@pytest.mark.parametrize("bad", ["zero", "nan"])
@pytest.mark.parametrize("col", [0, 5, 37])
def test_blocked_gj_unusable_pivot_reported_in_bounds(dgj, bad, col):
    rng = np.random.default_rng(6)
    n = 40
    A = rng.standard_normal((n, n))
    A[:, col] = 0.0 if bad == "zero" else np.nan
    status, piv, _ = run_blocked(dgj, [A], True)
    assert status == 8
    assert np.all((piv[0] >= 0) & (piv[0] < n))
