"""The C ABI library builds, loads and exports every entry point of include/fvc.h.
No compute calls: these run without a GPU."""
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "flowvc-1_amd")


def declared_symbols():
    text = open(os.path.join(REPO, "include", "fvc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fvc_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-s", "-C", PKG], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return os.path.join(PKG, "libfvc.so")


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ["fvc_create", "fvc_destroy", "fvc_load_slice", "fvc_advect", "fvc_locate", "fvc_ftle_frame",
              "fvc_ftle_output", "fvc_zsort", "fvc_strerror", "fvc_sync"]:
        assert s in syms


def test_library_exports_every_declared_symbol(built):
    out = subprocess.run(["nm", "-D", "--defined-only", built], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (fvc_[a-z0-9_]+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing


def test_library_loads_and_names_errors(built):
    import ctypes
    L = ctypes.CDLL(built)
    L.fvc_strerror.restype = ctypes.c_char_p
    assert L.fvc_strerror(-4) == b"Indeterminate neighbor"
    assert L.fvc_strerror(0) == b"ok"


def test_kernels_are_gfx950(built):
    """The shared object carries a gfx950 code object (no other targets)."""
    out = subprocess.run(["/opt/rocm/bin/roc-obj-ls", built], capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("roc-obj-ls unavailable")
    assert "gfx950" in out.stdout
