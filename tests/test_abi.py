"""The C-ABI library loads and exports every entry point that include/lettuce_hip.h
declares; host-side argument checking works without touching a GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lettuce_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(lt_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "lt_ctx_create" in syms and "lt_or_sweep" in syms and "lt_flow_run" in syms


def test_library_exports_every_declared_symbol():
    from lettuce_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    from lettuce_amd import _lib
    assert set(declared_symbols()) <= set(_lib.exported_symbols())


def test_library_is_gfx950_code_object():
    from lettuce_amd import _lib
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_null_arguments_are_rejected():
    from lettuce_amd import _lib
    lib = _lib.lib()
    assert lib.lt_ctx_create(0, None) == -1
    assert lib.lt_field_create(None, 4, 4, 4, 4, None) == -1
    assert b"null" in lib.lt_last_error()
    assert lib.lt_or_sweep(None, 1, 2) == -1
    assert lib.lt_observables(None, None) == -1
    assert lib.lt_ctx_read_rate(None, 1 << 21, None) == -1
    # destroying null handles is a no-op
    assert lib.lt_field_destroy(None) == 0
    assert lib.lt_ctx_destroy(None) == 0


def test_no_device_reports_error_not_crash():
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    from lettuce_amd import _lib
    lib = _lib.lib()
    h = ctypes.c_void_p()
    rc = lib.lt_ctx_create(0, ctypes.byref(h))
    assert rc < 0
    assert len(lib.lt_last_error()) > 0


def test_cpp_layer_compiles_and_links(tmp_path):
    """The header-only C++ layer builds a reference-style driver against the library."""
    import shutil
    import subprocess
    from lettuce_amd import _lib
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    exe = tmp_path / "driver"
    cmd = [gxx, "-std=c++17", "-O1", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(ROOT, "tests", "cpp", "driver_example.cpp"), "-o", str(exe),
           f"-L{os.path.dirname(_lib.LIB_PATH)}", "-llettuce_hip", "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
