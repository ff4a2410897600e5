"""The Gauss-Jordan inverses of kern-interp-old_amd/csrc/device_util.h run on
the CPU on emulated GPU threads (tests/cpp/simt/hip/hip_runtime.h), built
with AddressSanitizer and UBSan: every array the routines index (matrix,
pivot and permutation vectors, LDS scratch vectors) is a heap block of
exactly its kernel size, so an index past any of them fails the run.

They replace dgetrf + dgetri of the reference's root and X_rr LUs
(ki_mat.cpp:427-493, skel_factorization.cpp:98-128, :236-242) with the same
pivot order, so the checks are:

* the inverse against LAPACK's (numpy.linalg.inv) on well-conditioned
  matrices, and the pivot sequence against dgetrf's;
* the root matrix allskel_mat and the bordered matrix S of the annulus at
  id_tol 1e-13 (kern_interp_test.cpp:413-430) from the oracle: allskel_mat
  is numerically singular there (two pivots at roundoff level, cond ~4e17),
  and the routines must stay inside their arrays and give an S inverse to
  the accuracy cond(S) allows;
* the register-resident Gauss-Jordan of the one-wavefront elimination
  (wave_gj_reg: row i in lane i, no row interchanges, the pivot row
  broadcast through LDS) in FP64 and FP32 at every register width it has
  (16, 32, 64 rows), against LAPACK's inverse;
* a column without a usable pivot (all zero, or NaN): the routine reports
  it and every index stays in bounds.  With the first-candidate search of
  the earlier code (best index starting at n) a zero pivot turned the next
  column into NaN, the search returned n, and row n, piv[c] = n, sigma[n]
  and column n were read and written past the arrays.
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
PERM, BLOCK, WAVE, REG64, REG32 = 0, 1, 2, 3, 4


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


@pytest.mark.parametrize("variant,threads", [(PERM, 256), (BLOCK, 256), (WAVE, 64)])
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


def test_root_matrix_annulus_tol_1e13(emu, annulus_tight):
    # k_root's routine and output permutation, 1024 threads as launched
    A = annulus_tight.allskel_mat
    assert A.shape[0] == 217
    _, _, info = lapack.dgetrf(A)
    ok, piv, inv = run(emu, A, PERM, 1024)
    assert ok and np.all((piv >= 0) & (piv < A.shape[0]))
    assert np.isfinite(inv).all()


def test_bordered_matrix_annulus_tol_1e13(emu, annulus_tight):
    # k_dense_inverse<false>'s routine on S (n = 219), 1024 threads
    S = annulus_tight.S_mat
    assert S.shape[0] == 219
    ok, piv, inv = run(emu, S, BLOCK, 1024)
    assert ok
    ref = np.linalg.inv(S)
    bound = 1e-15 * np.linalg.cond(S, 1) * 100
    assert np.abs(inv - ref).max() <= bound * np.abs(ref).max()


@pytest.mark.parametrize("variant,threads", [(PERM, 256), (BLOCK, 256), (WAVE, 64)])
@pytest.mark.parametrize("bad", ["zero", "nan"])
def test_unusable_pivot_reported_in_bounds(emu, variant, threads, bad):
    rng = np.random.default_rng(5)
    n = 50
    A = rng.standard_normal((n, n))
    A[:, 5] = 0.0 if bad == "zero" else np.nan
    ok, piv, _ = run(emu, A, variant, threads)
    assert not ok
    assert np.all((piv >= 0) & (piv < n))


def test_largest_root_size(emu):
    # 1024 is the largest root / bordered system the library accepts
    rng = np.random.default_rng(3)
    n = 1024
    A = rng.standard_normal((n, n)) + 8 * np.eye(n)
    ok, piv, inv = run(emu, A, PERM, 256)
    assert ok
    assert np.abs(inv @ A - np.eye(n)).max() <= 1e-11


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
