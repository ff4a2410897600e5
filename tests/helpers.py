"""Shared test inputs: the workloads of the factor/solve benchmarks at small sizes."""
import numpy as np

from oracle import geometry as G
from oracle import kernel as K


def star_problem(n, src=(3.0, 3.0)):
    b = G.star(n)
    b.boundary_values = G.bc_point_source(b, src)
    return b


def exact_point_source(targets, src=(3.0, 3.0)):
    xy = np.asarray(targets).reshape(-1, 2)
    return np.log(np.sqrt((xy[:, 0] - src[0]) ** 2 + (xy[:, 1] - src[1]) ** 2)) / (2 * np.pi)


def gpu_tree(ctx, b, sd=1, leaf=64):
    return ctx.build_tree(b.points, b.normals, b.weights, b.curvatures, sd, leaf)


def pde_sd(pde):
    return 2 if pde == K.STOKES else 1
