"""The C-ABI library loads and exports every symbol include/shellgpu.h
declares (no GPU needed)."""
import ctypes
import os
import re

from shellgpu import HEADER_PATH, LIB_PATH, load_library


def declared_symbols():
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"\b(shellgpu_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB_PATH), "build with make -C a2d-shells_amd"
    lib = load_library()
    names = declared_symbols()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    raw = ctypes.CDLL(LIB_PATH)
    for n in names:
        getattr(raw, n)
