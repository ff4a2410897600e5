"""Factor + solve benchmark on one MI355X.

Workload: 2D Laplace interior Dirichlet problem, double-layer kernel on the star
r(theta) = 1 + 0.3 cos(5 theta), N = 65536 trapezoidal nodes, leaf 64, ID tol
1e-6, boundary data from a point source at (3, 3).  One step = quadtree
build + skeletonization + solve (+ FP64 iterative refinement for the FP32
path), from geometry and right-hand side resident in HBM to the density in
HBM.  Both arithmetic paths are timed: FP64, and FP32 elimination/solve with
FP64 ID and FP64 iterative refinement to relative residual 1e-10.

Prints one JSON line.  `value` is DOF/s = N / (seconds per step) of the
FP32+refinement path; the FP64 path's numbers sit beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

# the CPU baseline runs the oracle with single-threaded BLAS
os.environ["OPENBLAS_NUM_THREADS"] = "1"
os.environ.setdefault("MKL_NUM_THREADS", "1")

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "kern-interp-old_amd"))

import numpy as np  # noqa: E402

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--leaf", type=int, default=64)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the batched-instances and adjoint measurements")
    ap.add_argument("--cpu-n", type=int, default=65536)
    return ap.parse_args()


class DeviceProblem:
    """Geometry and right-hand sides resident in device memory (torch is only
    the allocator here; the library runs on its own HIP stream)."""

    def __init__(self, torch, geom, F):
        dev = torch.device("cuda:0")
        pts, nrm, w, kap = geom
        xy = pts.reshape(-1, 2)
        nm = nrm.reshape(-1, 2)
        self.arrays = [torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
                       for a in (xy[:, 0], xy[:, 1], nm[:, 0], nm[:, 1], w, kap)]
        self.F = torch.tensor(np.ascontiguousarray(F.T), dtype=torch.float64, device=dev)
        self.X = torch.zeros_like(self.F)
        self.n = w.size
        self.nrhs = F.shape[1]
        torch.cuda.synchronize()


def run_step(kigpu, ctx, prob, precision, leaf, tol):
    import ctypes as C
    lib = ctx.lib
    h = C.c_void_p()
    ptrs = [C.c_void_p(a.data_ptr()) for a in prob.arrays]
    kigpu._check(lib.kigpu_tree_build_device(ctx.h, *ptrs, prob.n, 1, leaf, C.byref(h)),
                 "kigpu_tree_build_device")
    opts = kigpu.FactorOpts(kigpu.LAPLACE, tol, precision, 16)
    kigpu._check(lib.kigpu_factor(ctx.h, h, C.byref(opts)), "kigpu_factor")
    ro = kigpu.RefineOpts(20 if precision == kigpu.FP32_REFINE else 0, 1e-10)
    info = kigpu.SolveInfo()
    kigpu._check(lib.kigpu_solve_device(ctx.h, h, C.c_void_p(prob.F.data_ptr()), prob.nrhs,
                                        C.c_void_p(prob.X.data_ptr()), None, C.byref(ro),
                                        C.byref(info)), "kigpu_solve_device")
    tim = ctx.timings()  # t_tree is the build's; factor and solve leave it alone
    return h, tim, {"iterations": info.iterations, "rel_residual": info.rel_residual}


def time_path(torch, kigpu, ctx, prob, precision, args):
    lib = ctx.lib
    for _ in range(args.warmup):
        h, _, _ = run_step(kigpu, ctx, prob, precision, args.leaf, args.tol)
        lib.kigpu_tree_destroy(h)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tims = []
    info = None
    h = None
    for i in range(args.steps):
        if h is not None:
            lib.kigpu_tree_destroy(h)
        h, tim, info = run_step(kigpu, ctx, prob, precision, args.leaf, args.tol)
        tims.append(tim)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    mu = prob.X.cpu().numpy().T.copy()
    return dt, tims, info, h, mu


def interior_error(kigpu, ctx, h, mu):
    from kigpu import workloads as W
    tgt = W.interior_disc_targets(1000, 1234)
    xy = tgt.reshape(-1, 2)
    ex = np.log(np.sqrt((xy[:, 0] - 3) ** 2 + (xy[:, 1] - 3) ** 2)) / (2 * np.pi)
    u = np.zeros(1000)
    m0 = np.ascontiguousarray(mu[:, 0])
    kigpu._check(ctx.lib.kigpu_forward_apply(ctx.h, h, kigpu.LAPLACE, tgt, 1000, m0, 1, u),
                 "kigpu_forward_apply")
    return float(np.linalg.norm(u - ex) / np.linalg.norm(ex))


def bench_batch(torch, kigpu, W, ctx, steps, warmup):
    """256 stars r = 1 + sum_{k=2..6} a_k cos(k t + phi_k), N ~ U{1024..16384}
    rounded to 64, Laplace, tol 1e-6, FP32 factor + FP64 refinement to 1e-10:
    one forest, one launch per tree level over every instance."""
    import ctypes as C
    sizes = W.batch_sizes(256, 4)
    geoms = [W.random_star(int(n), 4000 + i) for i, n in enumerate(sizes)]
    pts = np.concatenate([g[0] for g in geoms]).reshape(-1, 2)
    nrm = np.concatenate([g[1] for g in geoms]).reshape(-1, 2)
    w = np.concatenate([g[2] for g in geoms])
    kap = np.concatenate([g[3] for g in geoms])
    f = np.concatenate([W.point_source_bc(g[0]) for g in geoms])
    dev = torch.device("cuda:0")
    arrs = [torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
            for a in (pts[:, 0], pts[:, 1], nrm[:, 0], nrm[:, 1], w, kap)]
    dF = torch.tensor(f, dtype=torch.float64, device=dev)
    dX = torch.zeros_like(dF)
    counts = sizes.astype(np.int32)
    lib = ctx.lib

    def step():
        h = C.c_void_p()
        kigpu._check(lib.kigpu_tree_build_batch_device(
            ctx.h, len(counts), counts, *[C.c_void_p(a.data_ptr()) for a in arrs], 1, 64,
            C.byref(h)), "kigpu_tree_build_batch_device")
        opts = kigpu.FactorOpts(kigpu.LAPLACE, 1e-6, kigpu.FP32_REFINE, 16)
        kigpu._check(lib.kigpu_factor(ctx.h, h, C.byref(opts)), "kigpu_factor")
        ro = kigpu.RefineOpts(20, 1e-10)
        info = kigpu.SolveInfo()
        kigpu._check(lib.kigpu_solve_device(ctx.h, h, C.c_void_p(dF.data_ptr()), 1,
                                            C.c_void_p(dX.data_ptr()), None, C.byref(ro),
                                            C.byref(info)), "kigpu_solve_device")
        lib.kigpu_tree_destroy(h)
        return info

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        info = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    ndof = int(sizes.sum())
    return {"workload": "256 uneven random stars, N~U{1024..16384}/64, FP32+refine to 1e-10",
            "dofs": ndof, "s_per_step": dt, "dof_per_s": ndof / dt,
            "refine": {"iterations": info.iterations, "rel_residual": info.rel_residual}}


def bench_multi_rhs(torch, kigpu, W, ctx, args, counts=(16, 128)):
    """The same star workload with several right-hand sides (seeded point
    sources at radius 2-4 besides the one at (3, 3)): both paths, 1 warm-up
    and 2 timed steps each.  Interior error is that of the first column."""
    import argparse
    pts, nrm, w, kap = W.star(args.n)
    out = {}
    for nrhs in counts:
        srcs = [(3.0, 3.0)] + [tuple(s) for s in W.point_sources(nrhs - 1, 7)]
        F = np.stack([W.point_source_bc(pts, s) for s in srcs[:nrhs]], 1)
        prob = DeviceProblem(torch, (pts, nrm, w, kap), F)
        sub = argparse.Namespace(**{**vars(args), "steps": 2, "warmup": 1})
        leg = {}
        for name, prec in (("fp64", kigpu.FP64), ("fp32_refine", kigpu.FP32_REFINE)):
            dt, tims, info, h, mu = time_path(torch, kigpu, ctx, prob, prec, sub)
            err = interior_error(kigpu, ctx, h, mu)
            ctx.lib.kigpu_tree_destroy(h)
            tm = {k: float(np.median([t[k] for t in tims])) for k in tims[0]}
            leg[name] = {"s_per_step": dt, "rhs_dof_per_s": args.n * nrhs / dt,
                         "interior_rel_err_col0": err, "refine": info,
                         "t_solve": tm["t_solve"], "t_refine_apply": tm["t_refine_apply"]}
        out[str(nrhs)] = leg
        del prob
    return out


def bench_adjoint(kigpu, ctx, evals):
    """Shape gradient of L = sum u(x_j)^2 (1024 targets) in 16 Fourier
    coefficients of a star with N = 32768: factor, solve, transposed solve,
    sensitivities (FP32 factor refined to 1e-12)."""
    from kigpu import shape
    from kigpu import workloads as W
    c = 0.02 * (2 * W.uniform01(3, 16) - 1)
    tgt = W.interior_disc_targets(1024, 33, radius=0.3)
    star = shape.FourierStar(32768, 16)
    kw = dict(id_tol=1e-10, precision=kigpu.FP32_REFINE, refine_iters=10, rel_residual=1e-12)
    shape.loss_and_gradient(ctx, star, c, tgt, **kw)
    t0 = time.perf_counter()
    for _ in range(evals):
        L, g, info = shape.loss_and_gradient(ctx, star, c, tgt, **kw)
    dt = (time.perf_counter() - t0) / evals
    return {"workload": "star N=32768, 16 coefficients, 1024 targets, id_tol 1e-10",
            "s_per_gradient": dt, "loss": L, "grad_norm": float(np.linalg.norm(g)),
            "note": "host arrays in and out (includes PCIe copies)"}


def pmc_traffic(name):
    """HBM bytes per launch of a kernel from its committed rocprofv3 --pmc
    summary under profiles/ (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration), or None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(n, leaf, tol):
    """The oracle (numpy/scipy restatement, single-threaded BLAS) on the same workload."""
    from oracle import geometry as G
    from oracle import kernel as K
    from oracle import linear_solve as LS
    b = G.star(n)
    b.boundary_values = G.bc_point_source(b, (3.0, 3.0))
    t0 = time.perf_counter()
    tree, kern, fac = LS.factor(b, K.LAPLACE, tol, leaf)
    mu, _ = LS.linear_solve(fac, tree, b.boundary_values)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "DOF/s", "cores": 1, "kind": "port",
            "sample": f"one FP64 tree+factor+solve of the same star workload at N={n} "
                      f"(numpy/scipy oracle, OpenBLAS 1 thread): {dt:.2f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)
    import torch
    import kigpu
    from kigpu import workloads as W

    pts, nrm, w, kap = W.star(args.n)
    srcs = [(3.0, 3.0)] + [tuple(s) for s in W.point_sources(max(args.nrhs - 1, 0), 7)]
    F = np.stack([W.point_source_bc(pts, s) for s in srcs[: args.nrhs]], 1)
    ctx = kigpu.Context(0)
    prob = DeviceProblem(torch, (pts, nrm, w, kap), F)

    res = {}
    for name, prec in (("fp64", kigpu.FP64), ("fp32_refine", kigpu.FP32_REFINE)):
        dt, tims, info, h, mu = time_path(torch, kigpu, ctx, prob, prec, args)
        err = interior_error(kigpu, ctx, h, mu)
        ctx.lib.kigpu_tree_destroy(h)
        tm = {k: float(np.median([t[k] for t in tims])) for k in tims[0]}
        res[name] = {"s_per_step": dt, "dof_per_s": args.n / dt, "interior_rel_err": err,
                     "refine": info, "phases_s": {k: tm[k] for k in (
                         "t_tree", "t_lists", "t_id", "t_decouple", "t_root", "t_solve",
                         "t_refine_apply")}, "timings": tm}

    # roofline of the kernel that dominates the FP32+refinement step: the ID
    # kernel k_id<double, float> (ID matrix generation, pivoted QR to the rank
    # step and T = R11^-1 R12, in FP64), one launch per tree level, timed with
    # HIP events on the library's stream.  Algorithmic flops per box (m rows,
    # n columns, rank k, generation cost c_K per entry): c_K m n + 2 m n +
    # sum_{i<=k} [2 (m - i) + 4 (m - i)(n - i - 1)] + k^2 (n - k), summed over
    # the level's boxes.
    tf = res["fp32_refine"]["timings"]
    id_s = tf["id_kernel_s"] / max(tf["id_launches"], 1)
    id_f = tf["id_flops"] / max(tf["id_launches"], 1)
    id_achieved = id_f / id_s / 1e12 if id_s > 0 else 0.0
    roofline = {"bound": "mfma", "achieved": id_achieved, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": id_achieved / PEAK_FP64_TFLOPS,
                "traffic": pmc_traffic("pmc_id_laplace_star_N65536.md"),
                "kernel": "k_id<double,float> (ID matrix + pivoted QR + T in FP64, one launch "
                          "per level; each box is a chain of dependent Householder steps, "
                          "bound by latency, measured against the FP64 peak)",
                "avg_launch_us": id_s * 1e6, "launches_per_step": tf["id_launches"],
                "flops_per_launch": id_f}
    # the refinement's residual apply: the treecode (box series + direct near
    # field).  Algorithmic flops: 8 (P - 1) + 15 per (target, series box) for
    # the Horner evaluation of the P <= 48 terms the box needs for the
    # target's leaf (8 per term summed + 7 per pair), 10 per (target, source)
    # pair summed directly.
    tc_launches = max(tf["tc_launches"], 1)
    tc_eval_s = tf["tc_eval_s"] / tc_launches
    tc_flops = (8.0 * tf["tc_series_terms"] + 7.0 * tf["tc_series_pairs"]
                + 10.0 * tf["tc_direct_pairs"])
    refine_apply = {"kernel": "k_tc_coef + k_tc_eval (treecode, Laplace double layer)",
                    "applies_per_step": tf["tc_launches"],
                    "avg_apply_us": tf["tc_kernel_s"] / tc_launches * 1e6,
                    "avg_eval_us": tc_eval_s * 1e6,
                    "series_pairs": tf["tc_series_pairs"], "direct_pairs": tf["tc_direct_pairs"],
                    "series_terms": tf["tc_series_terms"],
                    "eval_achieved_tflops": tc_flops / tc_eval_s / 1e12 if tc_eval_s > 0 else 0.0,
                    "direct_sum_equivalent_pairs": float(args.n) ** 2}
    # the elimination kernel beside it (FP32 path)
    dec_s = tf["dec_kernel_s"] / max(tf["dec_launches"], 1)
    dec_f = tf["dec_flops"] / max(tf["dec_launches"], 1)
    dec_kernel = {"kernel": "k_decouple<float> (elimination and Schur update per box)",
                  "achieved": dec_f / dec_s / 1e12 if dec_s > 0 else 0.0,
                  "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "avg_launch_us": dec_s * 1e6,
                  "launches_per_step": tf["dec_launches"]}

    fast = res["fp32_refine"]
    out = {
        "metric": "factor+solve s & DOF/s at N=65536 FP32+refine vs FP64 GPU, rel err<=1e-6",
        "value": fast["dof_per_s"], "unit": "DOF/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": fast["s_per_step"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 elimination/solve + f64 ID, root LU and refinement",
        "data": "synthetic",
        "config": {"workload": f"laplace_star_N{args.n}_leaf{args.leaf}_tol{args.tol:g}",
                   "N": args.n, "nrhs": args.nrhs, "leaf": args.leaf, "id_tol": args.tol,
                   "refine_to": 1e-10},
        "fp64": {k: res["fp64"][k] for k in ("s_per_step", "dof_per_s", "interior_rel_err",
                                               "phases_s")},
        "fp32_refine": {k: fast[k] for k in ("s_per_step", "dof_per_s", "interior_rel_err",
                                             "refine", "phases_s")},
        "speedup_fp32_refine_vs_fp64": res["fp64"]["s_per_step"] / fast["s_per_step"],
        "roofline": roofline,
        "refine_apply": refine_apply,
        "dec_kernel": dec_kernel,
    }
    if not args.no_extra:
        out["multi_rhs"] = bench_multi_rhs(torch, kigpu, W, ctx, args)
        out["batch"] = bench_batch(torch, kigpu, W, ctx, 2, 1)
        out["adjoint"] = bench_adjoint(kigpu, ctx, 2)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.leaf, args.tol)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
