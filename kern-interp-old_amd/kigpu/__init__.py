"""ctypes binding of libkigpu.so (the C ABI declared in include/kigpu.h).

The Python side is plumbing: it moves numpy arrays across the C ABI.  All
computation happens in the HIP library; there is no CPU fallback, and
loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libkigpu.so")

LAPLACE, LAPLACE_NEUMANN, STOKES = 0, 1, 2
FP64, FP32_REFINE = 0, 1

_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class FactorOpts(C.Structure):
    _fields_ = [("pde", C.c_int), ("id_tol", C.c_double), ("precision", C.c_int),
                ("min_dofs", C.c_int)]


class RefineOpts(C.Structure):
    _fields_ = [("max_iters", C.c_int), ("rel_residual", C.c_double)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("rel_residual", C.c_double)]


class TreeInfo(C.Structure):
    _fields_ = [("num_levels", C.c_int), ("num_nodes", C.c_int), ("num_points", C.c_int),
                ("solution_dim", C.c_int), ("leaf_dofs", C.c_int), ("min", C.c_double),
                ("max", C.c_double)]


class Timings(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("t_tree", "t_lists", "t_id", "t_decouple", "t_root",
                                          "t_solve", "t_refine_apply")] + [
        ("syncs", C.c_int), ("id_launches", C.c_int), ("id_kernel_s", C.c_double),
        ("id_flops", C.c_double), ("dec_launches", C.c_int), ("dec_kernel_s", C.c_double),
        ("dec_flops", C.c_double), ("apply_launches", C.c_int), ("apply_kernel_s", C.c_double),
        ("apply_pairs", C.c_double), ("tc_launches", C.c_int), ("tc_kernel_s", C.c_double),
        ("tc_eval_s", C.c_double), ("tc_series_pairs", C.c_double),
        ("tc_direct_pairs", C.c_double), ("tc_series_terms", C.c_double),
        ("graph_replays", C.c_int), ("graph_captures", C.c_int)]


# every symbol include/kigpu.h declares
EXPORTS = [
    "kigpu_context_create", "kigpu_context_destroy", "kigpu_status_string", "kigpu_last_error",
    "kigpu_last_timings", "kigpu_tree_build", "kigpu_tree_build_device", "kigpu_tree_destroy", "kigpu_tree_get_info",
    "kigpu_tree_export_nodes", "kigpu_tree_export_boxes", "kigpu_tree_export_neighbors",
    "kigpu_set_border", "kigpu_factor", "kigpu_factor_ranks", "kigpu_factor_export_lists",
    "kigpu_factor_allskel", "kigpu_solve", "kigpu_solve_device", "kigpu_solve_device_noalloc",
    "kigpu_proxy_block", "kigpu_interp_decomp", "kigpu_schur_eliminate",
    "kigpu_domain_mask", "kigpu_domain_solution", "kigpu_tree_import", "kigpu_set_graphs",
    "kigpu_get_tuning", "kigpu_set_tuning", "kigpu_autotune",
    "kigpu_kernel_apply",
    "kigpu_kernel_apply_device", "kigpu_forward_apply", "kigpu_geometry_create",
    "kigpu_geometry_destroy", "kigpu_kernel_block", "kigpu_factor_export_active",
    "kigpu_solve_transpose", "kigpu_solve_transpose_device", "kigpu_kernel_apply_transpose",
    "kigpu_forward_apply_transpose", "kigpu_kernel_sensitivity", "kigpu_forward_sensitivity",
    "kigpu_tree_build_batch", "kigpu_tree_build_batch_device", "kigpu_tree_instances",
    "kigpu_kernel_apply_treecode", "kigpu_set_refine_apply", "kigpu_set_id_method",
]

_lib = None


def load() -> C.CDLL:
    """Load libkigpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                           "`python kern-interp-old_amd/build.py` (or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "kigpu_context_create": (C.c_int, [C.c_int, C.POINTER(P)]),
        "kigpu_context_destroy": (C.c_int, [P]),
        "kigpu_status_string": (C.c_char_p, [C.c_int]),
        "kigpu_last_error": (C.c_char_p, []),
        "kigpu_last_timings": (C.c_int, [P, C.POINTER(Timings)]),
        "kigpu_tree_build": (C.c_int, [P, _f64p, _f64p, _f64p, _f64p, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(P)]),
        "kigpu_tree_build_device": (C.c_int, [P, P, P, P, P, P, P, C.c_int, C.c_int, C.c_int,
                                              C.POINTER(P)]),
        "kigpu_tree_destroy": (C.c_int, [P]),
        "kigpu_tree_build_batch": (C.c_int, [P, C.c_int, _i32p, _f64p, _f64p, _f64p, _f64p,
                                             C.c_int, C.c_int, C.POINTER(P)]),
        "kigpu_tree_instances": (C.c_int, [P, C.POINTER(C.c_int), P, P, P]),
        "kigpu_tree_build_batch_device": (C.c_int, [P, C.c_int, _i32p, P, P, P, P, P, P, C.c_int,
                                                    C.c_int, C.POINTER(P)]),
        "kigpu_tree_get_info": (C.c_int, [P, C.POINTER(TreeInfo)]),
        "kigpu_tree_export_nodes": (C.c_int, [P, _i64p, _i32p, _i32p, _i32p, _f64p, _f64p]),
        "kigpu_tree_export_boxes": (C.c_int, [P, _i64p, P]),
        "kigpu_tree_export_neighbors": (C.c_int, [P, _i64p, P]),
        "kigpu_set_border": (C.c_int, [P, P, P, C.c_int]),
        "kigpu_factor": (C.c_int, [P, P, C.POINTER(FactorOpts)]),
        "kigpu_factor_ranks": (C.c_int, [P, _i32p, _i32p]),
        "kigpu_factor_export_lists": (C.c_int, [P, _i64p, P, _i64p, P]),
        "kigpu_factor_allskel": (C.c_int, [P, C.POINTER(C.c_int), P]),
        "kigpu_solve": (C.c_int, [P, P, _f64p, C.c_int, C.c_int, _f64p, C.c_int, P,
                                  C.POINTER(RefineOpts), C.POINTER(SolveInfo)]),
        "kigpu_solve_device": (C.c_int, [P, P, P, C.c_int, P, P, C.POINTER(RefineOpts),
                                         C.POINTER(SolveInfo)]),
        "kigpu_solve_device_noalloc": (C.c_int, [P, P, P, C.c_int, P]),
        "kigpu_geometry_create": (C.c_int, [P, _f64p, _f64p, _f64p, _f64p, C.c_int, P, C.c_int,
                                            C.POINTER(P)]),
        "kigpu_geometry_destroy": (C.c_int, [P]),
        "kigpu_kernel_block": (C.c_int, [P, P, C.c_int, C.c_int, _i32p, C.c_int, _i32p, C.c_int,
                                         C.c_int, _f64p]),
        "kigpu_proxy_block": (C.c_int, [P, P, C.c_int, _f64p, _f64p, C.c_int, C.c_double, _i32p,
                                        C.c_int, _f64p]),
        "kigpu_interp_decomp": (C.c_int, [P, _f64p, C.c_int, C.c_int, C.c_double,
                                          C.POINTER(C.c_int), _i32p, _f64p]),
        "kigpu_schur_eliminate": (C.c_int, [P, _f64p, C.c_int, C.c_int, _i32p, _f64p, _f64p,
                                            _f64p, _i32p, _f64p, _f64p, _f64p]),
        "kigpu_domain_mask": (C.c_int, [P, C.POINTER(Region), _f64p, C.c_int, P]),
        "kigpu_domain_solution": (C.c_int, [P, P, C.c_int, _f64p, C.c_int, _f64p, P, P, C.c_int,
                                            C.c_int, P, _f64p]),
        "kigpu_kernel_apply": (C.c_int, [P, P, C.c_int, _f64p, C.c_int, _f64p]),
        "kigpu_kernel_apply_treecode": (C.c_int, [P, P, _f64p, C.c_int, _f64p]),
        "kigpu_set_refine_apply": (C.c_int, [P, C.c_int]),
        "kigpu_set_id_method": (C.c_int, [P, C.c_int]),
        "kigpu_set_graphs": (C.c_int, [P, C.c_int]),
        "kigpu_get_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
        "kigpu_set_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
        "kigpu_autotune": (C.c_int, [P, P, C.POINTER(FactorOpts), C.c_int, C.POINTER(Tuning)]),
        "kigpu_kernel_apply_device": (C.c_int, [P, P, C.c_int, P, C.c_int, P]),
        "kigpu_forward_apply": (C.c_int, [P, P, C.c_int, _f64p, C.c_int, _f64p, C.c_int, _f64p]),
        "kigpu_solve_transpose": (C.c_int, [P, P, _f64p, C.c_int, C.c_int, _f64p, C.c_int,
                                            C.POINTER(RefineOpts), C.POINTER(SolveInfo)]),
        "kigpu_solve_transpose_device": (C.c_int, [P, P, P, C.c_int, P, C.POINTER(RefineOpts),
                                                   C.POINTER(SolveInfo)]),
        "kigpu_kernel_apply_transpose": (C.c_int, [P, P, C.c_int, _f64p, C.c_int, _f64p]),
        "kigpu_forward_apply_transpose": (C.c_int, [P, P, C.c_int, _f64p, C.c_int, _f64p,
                                                    C.c_int, _f64p]),
        "kigpu_kernel_sensitivity": (C.c_int, [P, P, C.c_int, _f64p, _f64p, _f64p]),
        "kigpu_forward_sensitivity": (C.c_int, [P, P, C.c_int, _f64p, C.c_int, _f64p, _f64p,
                                                _f64p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class Tuning(C.Structure):
    """kigpu_tuning: launch choices (see include/kigpu.h)."""
    _fields_ = [("version", C.c_int), ("dec_wave", C.c_int * 4), ("mid_on_side", C.c_int),
                ("pair_sweeps", C.c_int), ("apply_split_cap", C.c_int),
                ("fwd_split_cap", C.c_int)]

    def as_dict(self):
        return {"version": self.version, "dec_wave": list(self.dec_wave),
                "mid_on_side": self.mid_on_side, "pair_sweeps": self.pair_sweeps,
                "apply_split_cap": self.apply_split_cap, "fwd_split_cap": self.fwd_split_cap}

    @classmethod
    def from_dict(cls, d):
        t = cls()
        t.version = d["version"]
        for i, v in enumerate(d["dec_wave"]):
            t.dec_wave[i] = v
        for k in ("mid_on_side", "pair_sweeps", "apply_split_cap", "fwd_split_cap"):
            setattr(t, k, d[k])
        return t


class Region(C.Structure):
    """kigpu_region: Boundary::is_in_domain as data (kind DISC, WINDING or STAR)."""
    _fields_ = [("kind", C.c_int), ("eps", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
                ("radius", C.c_double), ("nholes", C.c_int), ("holes", C.c_void_p),
                ("nodes", C.c_void_p), ("num_outer", C.c_int), ("hole_nodes", C.c_void_p),
                ("nterms", C.c_int), ("amps", C.c_void_p), ("freqs", C.c_void_p),
                ("phases", C.c_void_p)]


REGION_DISC, REGION_WINDING, REGION_STAR = 0, 1, 2


def _keep(region, *arrays):
    region._arrays = arrays   # the structure holds raw pointers into these
    return region


def disc_region(center, radius, holes=(), eps=1e-2):
    """Circle / Annulus / Donut rule; holes: (cx, cy, r) triples."""
    h = np.ascontiguousarray(np.asarray(holes, np.float64).reshape(-1))
    r = Region(REGION_DISC, eps, float(center[0]), float(center[1]), float(radius),
               h.size // 3, h.ctypes.data if h.size else None, None, 0, None, 0, None, None, None)
    return _keep(r, h)


def winding_region(points, num_outer, hole_nodes=(), eps=1e-2):
    """CubicBoundary rule over the boundary points (outer section first)."""
    p = _f64(points).reshape(-1)
    hn = np.ascontiguousarray(hole_nodes, np.int32)
    r = Region(REGION_WINDING, eps, 0.0, 0.0, 0.0, hn.size, None, p.ctypes.data, int(num_outer),
               hn.ctypes.data if hn.size else None, 0, None, None, None)
    return _keep(r, p, hn)


def star_region(amps, freqs, phases, eps=1e-2):
    a, f, ph = (_f64(v).reshape(-1) for v in (amps, freqs, phases))
    r = Region(REGION_STAR, eps, 0.0, 0.0, 0.0, 0, None, None, 0, None, a.size, a.ctypes.data,
               f.ctypes.data, ph.ctypes.data)
    return _keep(r, a, f, ph)


class KigpuError(RuntimeError):
    def __init__(self, code, what):
        lib = load()
        msg = lib.kigpu_status_string(code).decode()
        detail = lib.kigpu_last_error().decode()
        super().__init__(f"{what}: {msg} ({detail})")
        self.code = code


def _check(code, what):
    if code != 0:
        raise KigpuError(code, what)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """A device + HIP stream (kigpu_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        _check(self.lib.kigpu_context_create(device, C.byref(h)), "kigpu_context_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.kigpu_context_destroy(self.h)
            self.h = None

    def set_refine_apply(self, mode: int):
        """Residual apply of the FP64 refinement: 0 treecode (default), 1 direct sum."""
        _check(self.lib.kigpu_set_refine_apply(self.h, mode), "kigpu_set_refine_apply")

    def set_id_method(self, method: int):
        """Box IDs: 0 by precision (FP32_REFINE Gram matrix, FP64 pivoted QR), 1 pivoted QR, 2 Gram."""
        _check(self.lib.kigpu_set_id_method(self.h, method), "kigpu_set_id_method")

    def tuning(self) -> dict:
        t = Tuning()
        _check(self.lib.kigpu_get_tuning(self.h, C.byref(t)), "kigpu_get_tuning")
        return t.as_dict()

    def set_tuning(self, d: dict):
        t = Tuning.from_dict(d)
        _check(self.lib.kigpu_set_tuning(self.h, C.byref(t)), "kigpu_set_tuning")

    def autotune(self, tree, pde=LAPLACE, id_tol=1e-6, precision=FP64, reps=3, min_dofs=16):
        """Measure the launch choices on `tree` (left factored with the
        winners) and keep them on this context; returns them."""
        o = FactorOpts(pde, id_tol, precision, min_dofs)
        t = Tuning()
        _check(self.lib.kigpu_autotune(self.h, tree.h, C.byref(o), reps, C.byref(t)),
               "kigpu_autotune")
        tree.pde = pde
        return t.as_dict()

    def set_graphs(self, on: bool):
        """Replay the solve sweeps from captured HIP graphs (repeated solves)."""
        _check(self.lib.kigpu_set_graphs(self.h, int(bool(on))), "kigpu_set_graphs")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def timings(self) -> dict:
        t = Timings()
        _check(self.lib.kigpu_last_timings(self.h, C.byref(t)), "kigpu_last_timings")
        return {n: getattr(t, n) for n, _ in Timings._fields_}

    def geometry(self, points, normals, weights, curvatures):
        return Geometry(self, points, normals, weights, curvatures)

    def interp_decomp(self, A, tol):
        """ki_Mat::id of a host matrix on the device: (k, perm, T); k = 0 means
        not compressible (perm, T None)."""
        A = np.asarray(A, np.float64)
        m, n = A.shape
        flat = np.ascontiguousarray(A.T).reshape(-1)   # column-major
        k = C.c_int(0)
        perm = np.zeros(n, np.int32)
        T = np.zeros(n * n, np.float64)
        _check(self.lib.kigpu_interp_decomp(self.h, flat, m, n, float(tol), C.byref(k), perm, T),
               "kigpu_interp_decomp")
        k = k.value
        if k == 0:
            return 0, None, None
        return k, perm.astype(np.int64), T[:k * (n - k)].reshape(n - k, k).T.copy()

    def schur_eliminate(self, K_BN, k, perm, T):
        """The arithmetic of SkelFactorization::decouple for one box:
        dict of X_rr, X_rr_lu, X_rr_piv (1-based), L, U, schur_update and the
        status code (0, or KIGPU_ERR_SINGULAR = 9)."""
        K_BN = np.asarray(K_BN, np.float64)
        n = K_BN.shape[0]
        r = n - k
        kf = np.ascontiguousarray(K_BN.T).reshape(-1)
        pm = np.ascontiguousarray(perm, np.int32)
        Tf = np.ascontiguousarray(np.asarray(T, np.float64).T).reshape(-1)
        out = {"X_rr": np.zeros(r * r), "X_rr_lu": np.zeros(r * r), "L": np.zeros(k * r),
               "U": np.zeros(r * k), "schur_update": np.zeros(k * k)}
        piv = np.zeros(max(r, 1), np.int32)
        st = self.lib.kigpu_schur_eliminate(self.h, kf, n, k, pm, Tf, out["X_rr"], out["X_rr_lu"],
                                            piv, out["L"], out["U"], out["schur_update"])
        if st not in (0, 9):
            _check(st, "kigpu_schur_eliminate")
        shapes = {"X_rr": (r, r), "X_rr_lu": (r, r), "L": (k, r), "U": (r, k),
                  "schur_update": (k, k)}
        res = {a: out[a].reshape(shapes[a][1], shapes[a][0]).T.copy() for a in out}
        res["X_rr_piv"] = piv[:r].copy()
        res["status"] = st
        return res

    def domain_mask(self, region, targets):
        """Boundary::is_in_domain of every target (2n interleaved or (n, 2))."""
        t = _f64(targets).reshape(-1)
        n = t.size // 2
        out = np.zeros(max(n, 1), np.uint8)
        _check(self.lib.kigpu_domain_mask(self.h, C.byref(region), t, n, out.ctypes.data),
               "kigpu_domain_mask")
        return out[:n].astype(bool)

    def build_tree(self, points, normals, weights, curvatures, sd=1, leaf_dofs=32):
        return Tree(self, points, normals, weights, curvatures, sd, leaf_dofs)

    def build_batch(self, boundaries, sd=1, leaf_dofs=32):
        """One forest over independent boundaries: a sequence of
        (points, normals, weights, curvatures) tuples, points/normals
        interleaved (2n,) or (n, 2)."""
        return Tree(self, None, None, None, None, sd, leaf_dofs, batch=boundaries)


class Geometry:
    """Boundary arrays on the device (kigpu_geom): the dense blocks
    Kernel::operator() and the proxy rows of Kernel::get_proxy_mat."""

    def __init__(self, ctx, points, normals, weights, curvatures):
        self.ctx, self.lib = ctx, ctx.lib
        p, nm = _f64(points).reshape(-1), _f64(normals).reshape(-1)
        w, c = _f64(weights).reshape(-1), _f64(curvatures).reshape(-1)
        self.npts = w.size
        h = C.c_void_p()
        _check(self.lib.kigpu_geometry_create(ctx.h, p, nm, w, c, self.npts, None, 0, C.byref(h)),
               "kigpu_geometry_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.kigpu_geometry_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def block(self, pde, tgt, src):
        sd = 2 if pde == STOKES else 1
        t = np.ascontiguousarray(tgt, np.int32)
        s_ = np.ascontiguousarray(src, np.int32)
        out = np.zeros(t.size * s_.size)
        _check(self.lib.kigpu_kernel_block(self.ctx.h, self.h, pde, sd, t, t.size, s_, s_.size, 0,
                                           out), "kigpu_kernel_block")
        return out.reshape(s_.size, t.size).T.copy()

    def proxy_block(self, pde, cx, cy, r, box, np_=128):
        """Rows [proxy <- box; box <- proxy] of Kernel::get_proxy_mat for the
        ring of np_ points of radius r about (cx, cy), weight 2 pi r / np_."""
        ang = [2 * np.pi * i * (1.0 / np_) for i in range(np_)]
        cs = np.array([math.cos(a) for a in ang])
        sn = np.array([math.sin(a) for a in ang])
        pp = np.ascontiguousarray(np.stack([cx + r * cs, cy + r * sn], 1).reshape(-1))
        pn = np.ascontiguousarray(np.stack([cs, sn], 1).reshape(-1))
        b = np.ascontiguousarray(box, np.int32)
        rows = (4 if pde == STOKES else 2) * np_
        out = np.zeros(rows * b.size)
        _check(self.lib.kigpu_proxy_block(self.ctx.h, self.h, pde, pp, pn, np_,
                                          2.0 * np.pi * r / np_, b, b.size, out),
               "kigpu_proxy_block")
        return out.reshape(b.size, rows).T.copy()


class Tree:
    """Geometry + quadtree + factorization handle (kigpu_tree)."""

    def __init__(self, ctx, points, normals, weights, curvatures, sd, leaf_dofs, batch=None):
        self.ctx = ctx
        self.lib = ctx.lib
        self.h = None
        self.sd = sd
        h = C.c_void_p()
        if batch is None:
            pts, nrm = _f64(points).reshape(-1), _f64(normals).reshape(-1)
            w, k = _f64(weights), _f64(curvatures)
            self.npts = w.size
            _check(self.lib.kigpu_tree_build(ctx.h, pts, nrm, w, k, self.npts, sd, leaf_dofs,
                                             C.byref(h)), "kigpu_tree_build")
        else:
            parts = [[_f64(a).reshape(-1) for a in b] for b in batch]
            counts = np.array([p[2].size for p in parts], np.int32)
            pts, nrm, w, k = (np.concatenate([p[i] for p in parts]) for i in range(4))
            self.npts = int(counts.sum())
            _check(self.lib.kigpu_tree_build_batch(ctx.h, len(parts), counts, pts, nrm, w, k, sd,
                                                   leaf_dofs, C.byref(h)),
                   "kigpu_tree_build_batch")
        self.ndof = sd * self.npts
        self.h = h
        self.pde = None

    def instances(self) -> dict:
        """Per-instance point offsets, root-box bounds and root skeleton sizes."""
        n = C.c_int()
        _check(self.lib.kigpu_tree_instances(self.h, C.byref(n), None, None, None),
               "kigpu_tree_instances")
        off = np.zeros(n.value + 1, np.int64)
        bounds = np.zeros(2 * n.value)
        rs = np.zeros(n.value, np.int32)
        _check(self.lib.kigpu_tree_instances(self.h, C.byref(n), off.ctypes.data_as(C.c_void_p),
                                             bounds.ctypes.data_as(C.c_void_p),
                                             rs.ctypes.data_as(C.c_void_p)),
               "kigpu_tree_instances")
        return {"point_offsets": off, "bounds": bounds.reshape(-1, 2), "root_skel": rs}

    def close(self):
        if self.h:
            self.lib.kigpu_tree_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        i = TreeInfo()
        _check(self.lib.kigpu_tree_get_info(self.h, C.byref(i)), "kigpu_tree_get_info")
        return {n: getattr(i, n) for n, _ in TreeInfo._fields_}

    def export_nodes(self) -> dict:
        inf = self.info()
        L, N = inf["num_levels"], inf["num_nodes"]
        out = {"level_offsets": np.zeros(L + 1, np.int64), "parent": np.zeros(N, np.int32),
               "children": np.zeros(4 * N, np.int32), "is_leaf": np.zeros(N, np.int32),
               "corners": np.zeros(8 * N), "side": np.zeros(N)}
        _check(self.lib.kigpu_tree_export_nodes(self.h, out["level_offsets"], out["parent"],
                                                out["children"], out["is_leaf"], out["corners"],
                                                out["side"]), "kigpu_tree_export_nodes")
        out["children"] = out["children"].reshape(N, 4)
        out["corners"] = out["corners"].reshape(N, 8)
        return out

    def _csr(self, fn, what):
        N = self.info()["num_nodes"]
        off = np.zeros(N + 1, np.int64)
        _check(fn(self.h, off, None), what)
        data = np.zeros(max(int(off[-1]), 1), np.int32)
        _check(fn(self.h, off, data.ctypes.data_as(C.c_void_p)), what)
        return off, data[: off[-1]]

    def export_boxes(self):
        return self._csr(self.lib.kigpu_tree_export_boxes, "kigpu_tree_export_boxes")

    def export_neighbors(self):
        return self._csr(self.lib.kigpu_tree_export_neighbors, "kigpu_tree_export_neighbors")

    def set_border(self, U, Psi):
        """QuadTree::U / ::Psi for holes: U (ndof x h), Psi (h x ndof)."""
        U = np.asfortranarray(U, dtype=np.float64)
        Psi = np.asfortranarray(Psi, dtype=np.float64)
        h = U.shape[1] if U.size else 0
        self.nborder = h
        if h == 0:
            _check(self.lib.kigpu_set_border(self.h, None, None, 0), "kigpu_set_border")
            return
        assert U.shape == (self.ndof, h) and Psi.shape == (h, self.ndof)
        self._U, self._Psi = U, Psi
        _check(self.lib.kigpu_set_border(self.h, U.ctypes.data_as(C.c_void_p),
                                         Psi.ctypes.data_as(C.c_void_p), h), "kigpu_set_border")

    def factor(self, pde=LAPLACE, id_tol=1e-6, precision=FP64, min_dofs=16):
        o = FactorOpts(pde, id_tol, precision, min_dofs)
        _check(self.lib.kigpu_factor(self.ctx.h, self.h, C.byref(o)), "kigpu_factor")
        self.pde = pde

    def ranks(self):
        N = self.info()["num_nodes"]
        r = np.zeros(N, np.int32)
        a = np.zeros(N, np.int32)
        _check(self.lib.kigpu_factor_ranks(self.h, r, a), "kigpu_factor_ranks")
        return r, a

    def lists(self):
        N = self.info()["num_nodes"]
        so = np.zeros(N + 1, np.int64)
        ro = np.zeros(N + 1, np.int64)
        _check(self.lib.kigpu_factor_export_lists(self.h, so, None, ro, None), "export_lists")
        s = np.zeros(max(int(so[-1]), 1), np.int32)
        r = np.zeros(max(int(ro[-1]), 1), np.int32)
        _check(self.lib.kigpu_factor_export_lists(self.h, so, s.ctypes.data_as(C.c_void_p), ro,
                                                  r.ctypes.data_as(C.c_void_p)), "export_lists")
        return so, s[: so[-1]], ro, r[: ro[-1]]

    def allskel(self):
        n = C.c_int()
        _check(self.lib.kigpu_factor_allskel(self.h, C.byref(n), None), "allskel")
        a = np.zeros(max(n.value, 1), np.int32)
        _check(self.lib.kigpu_factor_allskel(self.h, C.byref(n), a.ctypes.data_as(C.c_void_p)),
               "allskel")
        return a[: n.value]

    def solve(self, b, refine_iters=0, rel_residual=1e-10):
        b = np.asarray(b, np.float64)
        one = b.ndim == 1
        B = np.asfortranarray(b.reshape(self.ndof, -1))
        nrhs = B.shape[1]
        bflat = np.ascontiguousarray(B.T).reshape(-1)   # column-major
        x = np.zeros_like(bflat)
        h = getattr(self, "nborder", 0)
        alpha = np.zeros(max(h * nrhs, 1))
        ro = RefineOpts(refine_iters, rel_residual)
        info = SolveInfo()
        _check(self.lib.kigpu_solve(self.ctx.h, self.h, bflat, self.ndof, nrhs, x, self.ndof,
                                    alpha.ctypes.data_as(C.c_void_p) if h else None,
                                    C.byref(ro), C.byref(info)), "kigpu_solve")
        X = x.reshape(nrhs, self.ndof).T
        self.last_info = {"iterations": info.iterations, "rel_residual": info.rel_residual}
        mu = X[:, 0].copy() if one else np.array(X)
        if h:
            A = alpha[: h * nrhs].reshape(nrhs, h).T
            self.last_alpha = A[:, 0].copy() if one else np.array(A)
        return mu

    def solve_transpose(self, b, refine_iters=0, rel_residual=1e-10):
        """x = F^{-T} b (adjoint solve with the same factorization)."""
        b = np.asarray(b, np.float64)
        one = b.ndim == 1
        B = b.reshape(self.ndof, -1)
        nrhs = B.shape[1]
        bflat = np.ascontiguousarray(B.T).reshape(-1)
        x = np.zeros_like(bflat)
        ro = RefineOpts(refine_iters, rel_residual)
        info = SolveInfo()
        _check(self.lib.kigpu_solve_transpose(self.ctx.h, self.h, bflat, self.ndof, nrhs, x,
                                              self.ndof, C.byref(ro), C.byref(info)),
               "kigpu_solve_transpose")
        self.last_info = {"iterations": info.iterations, "rel_residual": info.rel_residual}
        X = x.reshape(nrhs, self.ndof).T
        return X[:, 0].copy() if one else np.array(X)

    def apply_transpose(self, lam, pde=None):
        pde = self.pde if pde is None else pde
        lam = np.asarray(lam, np.float64)
        one = lam.ndim == 1
        M = lam.reshape(self.ndof, -1)
        nrhs = M.shape[1]
        flat = np.ascontiguousarray(M.T).reshape(-1)
        y = np.zeros_like(flat)
        _check(self.lib.kigpu_kernel_apply_transpose(self.ctx.h, self.h, pde, flat, nrhs, y),
               "kigpu_kernel_apply_transpose")
        Y = y.reshape(nrhs, self.ndof).T
        return Y[:, 0].copy() if one else np.array(Y)

    def forward_transpose(self, targets, v, pde=None):
        pde = self.pde if pde is None else pde
        t = _f64(targets).reshape(-1)
        ntgt = t.size // 2
        sd = 2 if pde == STOKES else 1
        v = np.asarray(v, np.float64)
        one = v.ndim == 1
        V = v.reshape(sd * ntgt, -1)
        nrhs = V.shape[1]
        flat = np.ascontiguousarray(V.T).reshape(-1)
        y = np.zeros(self.ndof * nrhs)
        _check(self.lib.kigpu_forward_apply_transpose(self.ctx.h, self.h, pde, t, ntgt, flat,
                                                      nrhs, y), "kigpu_forward_apply_transpose")
        Y = y.reshape(nrhs, self.ndof).T
        return Y[:, 0].copy() if one else np.array(Y)

    def kernel_sensitivity(self, lam, mu):
        """(npts, 6): d(lam^T K mu)/d(x, y, nx, ny, w, kappa) per boundary point."""
        out = np.zeros(6 * self.ndof)
        _check(self.lib.kigpu_kernel_sensitivity(self.ctx.h, self.h, LAPLACE, _f64(lam),
                                                 _f64(mu), out), "kigpu_kernel_sensitivity")
        return out.reshape(-1, 6)

    def forward_sensitivity(self, targets, v, mu):
        """(npts, 6): d(v^T F mu)/d(x, y, nx, ny, w, kappa) per boundary point."""
        t = _f64(targets).reshape(-1)
        out = np.zeros(6 * self.ndof)
        _check(self.lib.kigpu_forward_sensitivity(self.ctx.h, self.h, LAPLACE, t, t.size // 2,
                                                  _f64(v), _f64(mu), out),
               "kigpu_forward_sensitivity")
        return out.reshape(-1, 6)

    def apply_treecode(self, mu):
        """K mu by the treecode (Laplace double layer; mu: (ndof,) or
        (ndof, nrhs)): the refinement's residual apply."""
        mu = np.asarray(mu, np.float64)
        one = mu.ndim == 1
        M = mu.reshape(self.ndof, -1)
        nrhs = M.shape[1]
        flat = np.ascontiguousarray(M.T).reshape(-1)
        y = np.zeros_like(flat)
        _check(self.lib.kigpu_kernel_apply_treecode(self.ctx.h, self.h, flat, nrhs, y),
               "kigpu_kernel_apply_treecode")
        Y = y.reshape(nrhs, self.ndof).T
        return Y[:, 0].copy() if one else np.array(Y)

    def apply(self, mu, pde=None):
        pde = self.pde if pde is None else pde
        mu = np.asarray(mu, np.float64)
        one = mu.ndim == 1
        M = mu.reshape(self.ndof, -1)
        nrhs = M.shape[1]
        flat = np.ascontiguousarray(M.T).reshape(-1)
        y = np.zeros_like(flat)
        _check(self.lib.kigpu_kernel_apply(self.ctx.h, self.h, pde, flat, nrhs, y),
               "kigpu_kernel_apply")
        Y = y.reshape(nrhs, self.ndof).T
        return Y[:, 0].copy() if one else np.array(Y)

    def domain_solution(self, targets, mu, alpha=None, hole_centers=None, region=None, pde=None):
        """boundary_integral_solve's domain values: forward(mu) + U_forward alpha,
        zero outside region (None: no mask)."""
        pde = self.pde if pde is None else pde
        t = _f64(targets).reshape(-1)
        ntgt = t.size // 2
        mu = np.asarray(mu, np.float64)
        one = mu.ndim == 1
        M = mu.reshape(self.ndof, -1)
        nrhs = M.shape[1]
        flat = np.ascontiguousarray(M.T).reshape(-1)
        nh = 0
        a_ptr = hc_ptr = None
        if alpha is not None:
            hc = _f64(hole_centers).reshape(-1)
            nh = hc.size // 2
            A = np.asarray(alpha, np.float64).reshape(-1, nrhs)
            af = np.ascontiguousarray(A.T).reshape(-1)
            a_ptr, hc_ptr = af.ctypes.data, hc.ctypes.data
        sd = 2 if pde == STOKES else 1
        u = np.zeros(sd * ntgt * nrhs)
        _check(self.lib.kigpu_domain_solution(self.ctx.h, self.h, pde, t, ntgt, flat, a_ptr,
                                              hc_ptr, nh, nrhs,
                                              C.addressof(region) if region is not None else None, u),
               "kigpu_domain_solution")
        U = u.reshape(nrhs, sd * ntgt).T
        return U[:, 0].copy() if one else np.array(U)

    def forward(self, targets, mu, pde=None):
        pde = self.pde if pde is None else pde
        t = _f64(targets).reshape(-1)
        ntgt = t.size // 2
        mu = np.asarray(mu, np.float64)
        one = mu.ndim == 1
        M = mu.reshape(self.ndof, -1)
        nrhs = M.shape[1]
        flat = np.ascontiguousarray(M.T).reshape(-1)
        sd = 2 if pde == STOKES else 1
        u = np.zeros(sd * ntgt * nrhs)
        _check(self.lib.kigpu_forward_apply(self.ctx.h, self.h, pde, t, ntgt, flat, nrhs, u),
               "kigpu_forward_apply")
        U = u.reshape(nrhs, sd * ntgt).T
        return U[:, 0].copy() if one else np.array(U)
