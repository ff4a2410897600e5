"""The C ABI library loads and exports every symbol include/kigpu.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kigpu.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:kigpu_status|const char\*)\s+(kigpu_\w+)\s*\(", txt, re.M)))


def test_header_lists_entry_points():
    names = declared()
    assert "kigpu_factor" in names and "kigpu_solve" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    import kigpu
    if not os.path.exists(kigpu.LIB_PATH):
        pytest.fail("libkigpu.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(kigpu.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(kigpu.EXPORTS) == declared()


def test_status_strings_without_gpu():
    import kigpu
    lib = kigpu.load()
    assert lib.kigpu_status_string(0) == b"ok"
    assert lib.kigpu_status_string(6) == b"no compression possible"
