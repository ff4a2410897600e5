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
DEFAULT_TUNING = {"version": 1, "dec_wave": [1, 1, 1, 1], "mid_on_side": 1, "pair_sweeps": 1,
                  "apply_split_cap": 64, "fwd_split_cap": 256}


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
    ap.add_argument("--tuning", default=os.path.join(ROOT, "kern-interp-old_amd", "tuning",
                                                     "gfx950.json"),
                    help="launch-tuning table to load (defaults when absent)")
    ap.add_argument("--autotune", default=None, metavar="PATH",
                    help="tune the launch choices on the workload first and write the table")
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


def pmc_traffic(name, key="traffic_bytes_per_launch"):
    """A figure from a committed rocprofv3 --pmc summary under profiles/:
    by default the HBM bytes per launch of a kernel (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration); None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


CPU_BASELINE_CODE = r"""
import json, os, sys, time
sys.path.insert(0, sys.argv[1])
from oracle import geometry as G
from oracle import kernel as K
from oracle import linear_solve as LS
n, leaf, tol, workers, budget = int(sys.argv[2]), int(sys.argv[3]), float(sys.argv[4]), int(sys.argv[5]), float(sys.argv[6])
b = G.star(n)
b.boundary_values = G.bc_point_source(b, (3.0, 3.0))
times = []
t_all = time.perf_counter()
while not times or (time.perf_counter() - t_all < budget and len(times) < 5):
    t0 = time.perf_counter()
    tree, kern, fac = LS.factor(b, K.LAPLACE, tol, leaf, fact_threads=workers)
    mu, _ = LS.linear_solve(fac, tree, b.boundary_values)
    times.append(time.perf_counter() - t0)
print(json.dumps({"times": times}))
"""


def cpu_baseline(n, leaf, tol, budget=12.0):
    """The oracle (numpy/scipy restatement of the reference path) on the same
    workload, its tree levels' boxes spread over the host cores as the
    reference spreads them over fact_threads OpenMP threads
    (skel_factorization.cpp:97-113): one forked process per core, BLAS
    single-threaded.  Runs in a child process (no GPU state), repeated
    until about `budget` seconds have passed (at most 5 runs)."""
    import subprocess
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    workers = max(1, min(16, ncpu))
    env = dict(os.environ, OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1", OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-c", CPU_BASELINE_CODE, ROOT, str(n), str(leaf), str(tol),
                        str(workers), str(budget)], capture_output=True, text=True, env=env,
                       timeout=600)
    if r.returncode != 0:
        raise RuntimeError("cpu baseline failed: " + r.stderr[-2000:])
    times = json.loads(r.stdout.strip().splitlines()[-1])["times"]
    dt = float(np.mean(times))
    return {"value": n / dt, "unit": "DOF/s", "cores": workers, "kind": "port",
            "sample": f"{len(times)} FP64 tree+factor+solve runs of the same star workload at "
                      f"N={n} (numpy/scipy oracle, each tree level's boxes over {workers} "
                      f"processes, OpenBLAS 1 thread each): {dt:.2f} s per run"}


def direct_residual(kigpu, ctx, h, mu, f):
    """||f - K mu|| / ||f|| with K applied by the FP64 direct sum (kigpu_kernel_apply)."""
    y = np.zeros_like(mu)
    kigpu._check(ctx.lib.kigpu_kernel_apply(ctx.h, h, kigpu.LAPLACE, np.ascontiguousarray(mu),
                                            1, y), "kigpu_kernel_apply")
    return float(np.linalg.norm(f - y) / np.linalg.norm(f))


def bench_stokes(kigpu, W, ctx, steps=2, warmup=1):
    """Stokes double layer in the unit disc with 8 seeded circular holes
    (16384 nodes outside, 2048 per hole: 65536 DOFs, 24 border unknowns),
    Stokeslet data from (2, 1): tree + factor + bordered solve per step, FP64
    and FP32 + refinement to 1e-10.  Host arrays in and out (the border
    matrices go to the device once per factorization)."""
    (pts, nrm, w, kap), holes = W.holed_disc(16384, 2048, 8, 7)
    f = W.stokeslet_bc(pts)
    U, Psi = W.border_matrices(True, holes, pts, w)
    out = {"workload": "stokes_holed_disc_16384+8x2048_leaf64_tol1e-06", "dofs": int(f.size),
           "border_unknowns": int(U.shape[1])}
    for name, prec in (("fp64", kigpu.FP64), ("fp32_refine", kigpu.FP32_REFINE)):
        refine = 20 if prec == kigpu.FP32_REFINE else 0

        def step():
            t = ctx.build_tree(pts, nrm, w, kap, 2, 64)
            t.set_border(U, Psi)
            t.factor(kigpu.STOKES, 1e-6, prec)
            mu = t.solve(f, refine_iters=refine, rel_residual=1e-10)
            return t, mu
        for _ in range(warmup):
            t, _ = step()
            t.close()
        t0 = time.perf_counter()
        for i in range(steps):
            if i:
                t.close()
            t, mu = step()
        dt = (time.perf_counter() - t0) / steps
        tim = ctx.timings()
        alpha = t.last_alpha
        r1 = t.apply(mu, kigpu.STOKES) + U @ alpha - f
        r2 = Psi @ mu - alpha
        bwd = float(np.sqrt(np.sum(r1 ** 2) + np.sum(r2 ** 2)) / np.linalg.norm(f))
        out[name] = {"s_per_step": dt, "dof_per_s": f.size / dt,
                     "direct_backward_error": bwd,
                     "refine": dict(t.last_info),
                     "phases_s": {k: tim[k] for k in ("t_tree", "t_lists", "t_id", "t_decouple",
                                                       "t_root", "t_solve", "t_refine_apply")}}
        t.close()
    out["speedup_fp32_refine_vs_fp64"] = out["fp64"]["s_per_step"] / out["fp32_refine"]["s_per_step"]
    return out


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

    from kigpu import autotune
    if args.autotune:
        for prec in (kigpu.FP64, kigpu.FP32_REFINE):
            t = ctx.build_tree(pts, nrm, w, kap, 1, args.leaf)
            autotune.tune(ctx, t, args.autotune, kigpu.LAPLACE, args.tol, prec, reps=3,
                          note=f"laplace star N={args.n} leaf {args.leaf}")
            t.close()
    table_path = args.autotune or args.tuning
    res = {}
    tunings = {}
    for name, prec in (("fp64", kigpu.FP64), ("fp32_refine", kigpu.FP32_REFINE)):
        ctx.set_tuning(DEFAULT_TUNING)
        entry = autotune.load(ctx, table_path, args.n, prec) if os.path.exists(table_path) else None
        tunings[name] = {"table": os.path.relpath(table_path, ROOT) if entry else None,
                         **ctx.tuning()}
        dt, tims, info, h, mu = time_path(torch, kigpu, ctx, prob, prec, args)
        err = interior_error(kigpu, ctx, h, mu)
        dres = direct_residual(kigpu, ctx, h, mu[:, 0], F[:, 0])
        ctx.lib.kigpu_tree_destroy(h)
        tm = {k: float(np.median([t[k] for t in tims])) for k in tims[0]}
        res[name] = {"s_per_step": dt, "dof_per_s": args.n / dt, "interior_rel_err": err,
                     "direct_rel_residual": dres,
                     "refine": info, "phases_s": {k: tm[k] for k in (
                         "t_tree", "t_lists", "t_id", "t_decouple", "t_root", "t_solve",
                         "t_refine_apply")}, "timings": tm}

    # roofline of the elimination kernel k_decouple_wave<float> (with the
    # workgroup kernel for boxes above 64 DOFs): one wavefront per box, the
    # Schur-update products on the FP32 matrix cores, timed per tree level
    # with HIP events on the library's stream around both streams' launches.
    # Algorithmic flops per box (n active DOFs, rank k, r = n - k): c_K n^2
    # for the K_BN entries + 6 k^2 r + 8 k r^2 + 2/3 r^3, summed over the
    # level; measured against the FP32 matrix peak.
    tf = res["fp32_refine"]["timings"]
    dec_s = tf["dec_kernel_s"] / max(tf["dec_launches"], 1)
    dec_f = tf["dec_flops"] / max(tf["dec_launches"], 1)
    dec_achieved = dec_f / dec_s / 1e12 if dec_s > 0 else 0.0
    waves = tunings["fp32_refine"]["dec_wave"]
    kernels = []
    if 0 in waves:
        kernels.append("k_decouple<float> (one 1024-thread workgroup per box, products as dot "
                       "products over LDS, block Gauss-Jordan)")
    if 1 in waves:
        kernels.append("k_decouple_wave<float> (one wavefront per box, products on "
                       "v_mfma_f32_16x16x4f32, register Gauss-Jordan)")
    roofline = {"bound": "mfma", "achieved": dec_achieved, "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": dec_achieved / PEAK_FP32_TFLOPS,
                "traffic": pmc_traffic("pmc_decouple_laplace_star_N65536.md"),
                "kernel": "elimination and Schur update: " + "; ".join(kernels) +
                          f" (per box-size class as tuned: dec_wave {waves})",
                "limit": "latency of each box's dependent Gauss-Jordan steps; a level has few "
                         "boxes per CU",
                "avg_launch_us": dec_s * 1e6, "launches_per_step": tf["dec_launches"],
                "flops_per_launch": dec_f}
    # the ID kernels beside it (one level's ID launches, timed together):
    # the Gram-matrix ID for boxes of up to 64 DOFs (A^T A on the FP64 matrix
    # cores, pivoted Cholesky) and the pivoted Householder QR for the larger
    # ones, measured against the FP64 peak (vector and matrix alike)
    id_s = tf["id_kernel_s"] / max(tf["id_launches"], 1)
    id_f = tf["id_flops"] / max(tf["id_launches"], 1)
    id_kernel = {"kernel": "k_id_gram<float> (ID matrix generated into A^T A on "
                           "v_mfma_f64_16x16x4f64, pivoted Cholesky, T) and k_id<double,float> "
                           "(ID matrix + pivoted Householder QR + T in FP64, boxes above 64 DOFs)",
                 "limit": "latency of the dependent pivot steps; few boxes per launch at the "
                          "coarse levels",
                 "achieved": id_f / id_s / 1e12 if id_s > 0 else 0.0,
                 "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                 "avg_launch_us": id_s * 1e6, "launches_per_step": tf["id_launches"],
                 "traffic": pmc_traffic("pmc_id_laplace_star_N65536.md"),
                 "mfma_util_percent": pmc_traffic("pmc_mfma_laplace_star_N65536.md",
                                                  "mfma_util_id_percent")}
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
    fast = res["fp32_refine"]
    out = {
        "metric": "factor+solve s & DOF/s at N=65536 FP32+refine vs FP64 GPU, rel err<=1e-6",
        "value": fast["dof_per_s"], "unit": "DOF/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": fast["s_per_step"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 elimination (matrix cores) and solve + f64 ID, root LU and refinement",
        "data": "synthetic",
        "config": {"workload": f"laplace_star_N{args.n}_leaf{args.leaf}_tol{args.tol:g}",
                   "N": args.n, "nrhs": args.nrhs, "leaf": args.leaf, "id_tol": args.tol,
                   "refine_to": 1e-10},
        "fp64": {k: res["fp64"][k] for k in ("s_per_step", "dof_per_s", "interior_rel_err",
                                               "direct_rel_residual", "phases_s")},
        "fp32_refine": {k: fast[k] for k in ("s_per_step", "dof_per_s", "interior_rel_err",
                                             "direct_rel_residual", "refine", "phases_s")},
        "speedup_fp32_refine_vs_fp64": res["fp64"]["s_per_step"] / fast["s_per_step"],
        "tuning": tunings,
        "roofline": roofline,
        "id_kernel": id_kernel,
        "refine_apply": refine_apply,
    }
    if not args.no_extra:
        out["multi_rhs"] = bench_multi_rhs(torch, kigpu, W, ctx, args)
        out["batch"] = bench_batch(torch, kigpu, W, ctx, 2, 1)
        out["adjoint"] = bench_adjoint(kigpu, ctx, 2)
        out["stokes_holes"] = bench_stokes(kigpu, W, ctx)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.leaf, args.tol)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
