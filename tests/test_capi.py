"""The C ABI library loads and exports every entry point include/k2eval.h declares (no GPU
calls: this runs on machines without a GPU)."""
import ctypes
import os
import re
import subprocess

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "k2eval.h")
LIB = os.path.join(ROOT, "mertens_bound_amd", "libk2eval.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("if", "sizeof")))


def _ensure_built():
    if not os.path.exists(LIB):
        from mertens_bound_amd import _build
        _build.build()


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for must in ("k2_table_create", "k2_eval", "k2_eval_device", "k2_eval_batched_device",
                 "k2_polish_device", "k2_phase_words", "compute_peak_AB",
                 "polish_peaks_with_weights", "k2_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    _ensure_built()
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    assert set(declared_functions()) <= exported


def test_library_has_gfx950_code_object():
    _ensure_built()
    out = subprocess.run(["/opt/rocm/bin/roc-obj-ls", LIB], capture_output=True, text=True)
    if out.returncode != 0:
        data = open(LIB, "rb").read()
        assert b"gfx950" in data
    else:
        assert "gfx950" in out.stdout


def test_error_path_without_device_work():
    _ensure_built()
    lib = ctypes.CDLL(LIB)
    lib.k2_last_error.restype = ctypes.c_char_p
    lib.k2_table_create.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p]
    # a null output handle is rejected before any HIP call
    assert lib.k2_table_create(None, 10, None, None, None) == 1
    assert b"null" in lib.k2_last_error()


def test_header_is_plain_c_and_cxx():
    """The boundary header compiles on its own as strict C99 and as C++17 (a C caller, cgo or a
    ctypes user's C shim includes it without HIP or torch headers)."""
    import shutil
    import pytest
    for cc, args in (("gcc", ["-std=c99", "-pedantic", "-x", "c"]),
                     ("g++", ["-std=c++17", "-x", "c++"])):
        if shutil.which(cc) is None:
            pytest.skip(f"{cc} not available")
        r = subprocess.run([cc, *args, "-Wall", "-Wextra", "-Werror", "-fsyntax-only", HEADER],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
