"""The C ABI: libzk.so loads and exports every entry point include/zk.h
declares; without a GPU the constructor fails loudly (no CPU fallback)."""
import re
import subprocess
from pathlib import Path

import pytest

from zuikakucfd_amd import zk

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    txt = (ROOT / "include" / "zk.h").read_text()
    return sorted(set(re.findall(r"\b(zk_[a-z_]+)\s*\(", txt)))


def test_header_declares_python_binding_set():
    assert declared_symbols() == sorted(zk.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = zk.load_library()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(zk.LIB_PATH)], capture_output=True, text=True).stdout
    for name in declared_symbols():
        assert re.search(rf"\bT {name}$", out, re.M), name


def test_library_carries_gfx950_code_object():
    blob = zk.LIB_PATH.read_bytes()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_create_rejects_bad_config():
    lib = zk.load_library()
    import ctypes as C

    cfg = zk.ZkConfig(zk.METHOD_DG, 7, 0, 0, 1, 50.0, 0)
    h = C.c_void_p()
    assert lib.zk_create(C.byref(cfg), C.byref(h)) != 0
    assert b"P must be" in lib.zk_last_error(None)


def test_no_gpu_fails_loudly():
    """In a process with no GPU the solver raises instead of falling back."""
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except Exception:
        pass
    with pytest.raises(zk.ZkError):
        zk.Solver(method=zk.METHOD_WENO)
