"""kigpu_solve_device_noalloc against kigpu_solve_device (refine max_iters =
0) through the C ABI: the same kernels in the same order, so the solutions
must agree bit for bit.  Two calls in a row on each factor, the second with
fewer right-hand sides, so the second reads the static scratch the first
left behind; an FP32 factor (scratch staging) and an FP64 factor (in place).
Device buffers come from torch (only the allocator here), imported before
the library is loaded so that both use one HIP runtime."""
import ctypes as C

import numpy as np
import pytest
import torch

import kigpu
from kigpu import workloads as W


def dev_solve(torch, ctx, tree, F, noalloc):
    dev = torch.device("cuda:0")
    nrhs = F.shape[1]
    dF = torch.tensor(np.ascontiguousarray(F.T), dtype=torch.float64, device=dev)
    dX = torch.zeros_like(dF)
    torch.cuda.synchronize()
    lib = ctx.lib
    if noalloc:
        st = lib.kigpu_solve_device_noalloc(ctx.h, tree.h, C.c_void_p(dF.data_ptr()), nrhs,
                                            C.c_void_p(dX.data_ptr()))
    else:
        ro = kigpu.RefineOpts(0, 1e-10)
        info = kigpu.SolveInfo()
        st = lib.kigpu_solve_device(ctx.h, tree.h, C.c_void_p(dF.data_ptr()), nrhs,
                                    C.c_void_p(dX.data_ptr()), None, C.byref(ro),
                                    C.byref(info))
    kigpu._check(st, "solve")
    return dX.cpu().numpy().T.copy()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [kigpu.FP32_REFINE, kigpu.FP64])
def test_noalloc_matches_solve_device(ctx, precision):
    print(f"[noalloc] star N=16384 precision {precision}", flush=True)
    n = 16384
    pts, nrm, w, kap = W.star(n)
    tree = ctx.build_tree(pts, nrm, w, kap, 1, 64)
    tree.factor(kigpu.LAPLACE, 1e-6, precision)
    for nrhs in (8, 3):   # second call smaller than the first
        src = W.point_sources(nrhs, 21 + nrhs)
        F = np.stack([W.point_source_bc(pts, s) for s in src], 1)
        ref = dev_solve(torch, ctx, tree, F, noalloc=False)
        got = dev_solve(torch, ctx, tree, F, noalloc=True)
        assert np.array_equal(got, ref), np.abs(got - ref).max()
    tree.close()


@pytest.mark.gpu
def test_noalloc_rejects_oversize(ctx):
    print("[noalloc] oversize request is rejected", flush=True)
    n = 4096
    pts, nrm, w, kap = W.star(n)
    tree = ctx.build_tree(pts, nrm, w, kap, 1, 64)
    tree.factor(kigpu.LAPLACE, 1e-6, kigpu.FP32_REFINE)
    nrhs = (1 << 22) // n + 1   # ndof * nrhs just past the scratch
    dev = torch.device("cuda:0")
    dF = torch.zeros((nrhs, n), dtype=torch.float64, device=dev)
    dX = torch.zeros_like(dF)
    torch.cuda.synchronize()
    st = ctx.lib.kigpu_solve_device_noalloc(ctx.h, tree.h, C.c_void_p(dF.data_ptr()), nrhs,
                                            C.c_void_p(dX.data_ptr()))
    assert st == 1   # KIGPU_ERR_ARG, before any launch
    tree.close()
