"""kigpu_autotune and the tuning table on the GPU.

The tuner factors a star N=16384 (leaf 64) per elimination choice and keeps
the fastest per box-size class; the winners must be valid choices, the tree
must be left factored and solve to the refinement target, and a table
written by one context must set the same choices in another.  The launch
choices change only how the work is cut: with every box on the
one-workgroup elimination and unpaired sweeps the FP64 density agrees with
the default choices (one wavefront per box) to 1e-10 relative.
"""
import numpy as np
import pytest

import kigpu
from kigpu import autotune
from oracle import kernel as K

from helpers import gpu_tree, star_problem


@pytest.mark.gpu
def test_autotune_and_table(tmp_path):
    print("[autotune] star 16384 FP32+refine", flush=True)
    b = star_problem(16384)
    ctx = kigpu.Context(0)
    try:
        t = gpu_tree(ctx, b, 1, 64)
        path = str(tmp_path / "tuning.json")
        e = autotune.tune(ctx, t, path, K.LAPLACE, 1e-6, kigpu.FP32_REFINE, reps=2)
        tu = e["tuning"]
        print("  tuned:", tu, flush=True)
        assert tu["version"] == 1
        assert all(v in (0, 1) for v in tu["dec_wave"])
        assert tu["mid_on_side"] in (0, 1) and tu["pair_sweeps"] in (0, 1)
        assert ctx.tuning() == tu
        # left factored: solve to the refinement target
        mu = t.solve(b.boundary_values, refine_iters=10, rel_residual=1e-10)
        r = t.apply(mu) - b.boundary_values
        assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(b.boundary_values)
    finally:
        ctx.close()
    ctx2 = kigpu.Context(0)
    try:
        got = autotune.load(ctx2, path, 16384, kigpu.FP32_REFINE)
        assert got is not None and ctx2.tuning() == tu
    finally:
        ctx2.close()


@pytest.mark.gpu
@pytest.mark.parametrize("dec", [[0, 0, 0, 0], [1, 0, 0, 0]])
def test_choices_change_only_rounding(ctx, dec):
    # the default one-wave elimination against the one-workgroup kernel and
    # a mix of the two (with unpaired sweeps and other split caps)
    print(f"[autotune] star 16384 FP64: elimination choice {dec}, unpaired sweeps", flush=True)
    b = star_problem(16384)
    t = gpu_tree(ctx, b, 1, 64)
    t.factor(K.LAPLACE, 1e-6, kigpu.FP64)
    mu0 = t.solve(b.boundary_values)
    alt = kigpu.Context(0)
    try:
        alt.set_tuning({"version": 1, "dec_wave": dec, "mid_on_side": 0,
                        "pair_sweeps": 0, "apply_split_cap": 16, "fwd_split_cap": 32})
        t2 = gpu_tree(alt, b, 1, 64)
        t2.factor(K.LAPLACE, 1e-6, kigpu.FP64)
        mu1 = t2.solve(b.boundary_values)
        assert np.linalg.norm(mu1 - mu0) <= 1e-10 * np.linalg.norm(mu0)
        tgt = np.array([[0.1, 0.2], [-0.3, 0.4], [0.5, -0.2]])
        u0, u1 = t.forward(tgt, mu0), t2.forward(tgt, mu0)
        assert np.linalg.norm(u1 - u0) <= 1e-13 * np.linalg.norm(u0)
        with pytest.raises(kigpu.KigpuError):
            alt.set_tuning({"version": 2, "dec_wave": [1, 1, 1, 1], "mid_on_side": 1,
                            "pair_sweeps": 1, "apply_split_cap": 64, "fwd_split_cap": 256})
        for bad in (2, 3, -1):
            with pytest.raises(kigpu.KigpuError):
                alt.set_tuning({"version": 1, "dec_wave": [bad, 1, 1, 1], "mid_on_side": 1,
                                "pair_sweeps": 1, "apply_split_cap": 64, "fwd_split_cap": 256})
    finally:
        alt.close()

