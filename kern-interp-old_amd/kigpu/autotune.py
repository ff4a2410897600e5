"""Launch-tuning table: measured kigpu_tuning choices kept in a versioned JSON
file, keyed by (GPU architecture, arithmetic, size bucket = round(log2 DOFs)).

    from kigpu import autotune
    entry = autotune.tune(ctx, tree, path, pde=..., id_tol=..., precision=...)
    autotune.load(ctx, path, ndof, precision)     # later, in another process

`tune` runs kigpu_autotune on a representative tree (left factored with the
winners) and merges the result into the file; `load` sets a context's tuning
from the entry of the nearest size bucket of the same architecture and
arithmetic, or leaves the defaults when the file has none.
"""
from __future__ import annotations

import json
import math
import os

from . import FP32_REFINE

TABLE_VERSION = 1
ARCH = "gfx950"   # the library is built for this architecture only


def bucket(ndof: int) -> int:
    return int(round(math.log2(max(int(ndof), 1))))


def key(arch: str, precision: int, ndof: int) -> str:
    return f"{arch}/{'fp32' if precision == FP32_REFINE else 'fp64'}/{bucket(ndof)}"


def read(path: str) -> dict:
    if not os.path.exists(path):
        return {"version": TABLE_VERSION, "entries": {}}
    with open(path) as f:
        table = json.load(f)
    if table.get("version") != TABLE_VERSION:
        raise ValueError(f"{path}: tuning table version {table.get('version')}, "
                         f"this library reads {TABLE_VERSION}")
    return table


def write(path: str, table: dict) -> None:
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune(ctx, tree, path: str | None, pde: int, id_tol: float, precision: int, reps: int = 3,
         arch: str = ARCH, note: str = "") -> dict:
    tuning = ctx.autotune(tree, pde=pde, id_tol=id_tol, precision=precision, reps=reps)
    ndof = tree.ndof
    entry = {"tuning": tuning, "ndof": ndof, "pde": pde, "id_tol": id_tol, "reps": reps,
             "note": note}
    if path:
        table = read(path)
        table["entries"][key(arch, precision, ndof)] = entry
        write(path, table)
    return entry


def lookup(table: dict, ndof: int, precision: int, arch: str = ARCH):
    want = key(arch, precision, ndof).rsplit("/", 1)[0] + "/"
    cands = [(abs(int(k.rsplit("/", 1)[1]) - bucket(ndof)), k)
             for k in table["entries"] if k.startswith(want)]
    if not cands:
        return None
    return table["entries"][min(cands)[1]]


def load(ctx, path: str, ndof: int, precision: int, arch: str = ARCH):
    """Set ctx's tuning from the table; returns the entry used, or None."""
    entry = lookup(read(path), ndof, precision, arch)
    if entry is not None:
        ctx.set_tuning(entry["tuning"])
    return entry
