"""The C-ABI library loads without a GPU and exports every entry point that
include/stwo_rocm.h declares; the Python binding binds exactly that set."""
import ctypes
import os
import re
import subprocess

import stwo_rocm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "stwo_rocm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(stwo_rocm_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_declared_symbols():
    lib = stwo_rocm.load_library()
    decl = declared_symbols()
    assert len(decl) >= 30
    for name in decl:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", stwo_rocm.LIB_PATH], capture_output=True, text=True).stdout
    exported = sorted(set(re.findall(r" T (stwo_rocm_[a-z0-9_]+)", out)))
    assert exported == decl


def test_python_binding_covers_header():
    assert sorted(stwo_rocm.SYMBOLS) == declared_symbols()


def test_library_is_gfx950_code():
    # the fat binary embeds code objects for exactly one target: gfx950
    data = open(stwo_rocm.LIB_PATH, "rb").read()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx[0-9a-z]+)", data))
    assert targets == {b"gfx950"}


def test_context_creation_fails_loudly_without_device():
    # On a machine without a HIP device the binding must raise, never fall back.
    import pytest

    if os.path.exists("/dev/kfd"):
        pytest.skip("a HIP device node exists here")
    with pytest.raises(stwo_rocm.RocmError):
        stwo_rocm.Context(0)
