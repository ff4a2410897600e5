"""The launch-tuning table (kigpu/autotune.py) on the CPU: keys, nearest
size-bucket lookup, versioned read/write round trip."""
import json

import pytest

from kigpu import FP32_REFINE, FP64, autotune


def entry(mark):
    return {"tuning": {"version": 1, "dec_wave": [1, 1, 0, mark], "mid_on_side": 1,
                       "pair_sweeps": 1, "apply_split_cap": 64, "fwd_split_cap": 256},
            "ndof": 0, "pde": 0, "id_tol": 1e-6, "reps": 3, "note": ""}


def test_keys_and_buckets():
    assert autotune.bucket(65536) == 16 and autotune.bucket(90000) == 16
    assert autotune.bucket(100000) == 17
    assert autotune.key("gfx950", FP32_REFINE, 65536) == "gfx950/fp32/16"
    assert autotune.key("gfx950", FP64, 4096) == "gfx950/fp64/12"


def test_nearest_bucket_lookup():
    table = {"version": 1, "entries": {"gfx950/fp32/12": entry(0), "gfx950/fp32/16": entry(1),
                                       "gfx950/fp64/16": entry(0), "gfx942/fp32/14": entry(0)}}
    assert autotune.lookup(table, 65536, FP32_REFINE)["tuning"]["dec_wave"][3] == 1
    assert autotune.lookup(table, 16384, FP32_REFINE)["tuning"]["dec_wave"][3] == 0   # 14 -> 12 or 16
    assert autotune.lookup(table, 1 << 20, FP32_REFINE)["tuning"]["dec_wave"][3] == 1
    assert autotune.lookup(table, 4096, FP64)["tuning"]["dec_wave"][3] == 0
    assert autotune.lookup(table, 4096, FP64, arch="gfx90a") is None


def test_round_trip_and_version(tmp_path):
    p = str(tmp_path / "t.json")
    t = autotune.read(p)
    assert t == {"version": 1, "entries": {}}
    t["entries"]["gfx950/fp32/16"] = entry(1)
    autotune.write(p, t)
    assert autotune.read(p) == t
    bad = dict(t, version=99)
    with open(p, "w") as f:
        json.dump(bad, f)
    with pytest.raises(ValueError):
        autotune.read(p)
