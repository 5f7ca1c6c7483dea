"""The C-ABI library loads and exports every function include/lhgpu.h
declares (no device calls: runs without a GPU)."""
import ctypes
import os
import re

from rnl_multiblock_nofft_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(lh_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == names


def test_error_path_without_context():
    L = _lib.load()
    rc = L.lh_ctx_create(None, 0, None)
    assert rc == -1
    assert b"bad arguments" in L.lh_last_error()
