"""The C-ABI library loads and exports every function include/*.h declares;
the C API conformance program compiles against include/xtb.h (CPU only)."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "xtb_amd", "libxtb_amd.so")


def declared_functions():
    names = set()
    for h in ("xtb.h", "xtb_amd_batch.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
        for m in re.finditer(r"\b(xtb(?:amd)?_[A-Za-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    names.discard("xtb_delete")
    return sorted(names)


def test_header_parse():
    names = declared_functions()
    assert "xtb_singlepoint" in names and "xtbamd_batch_new" in names
    assert len(names) >= 50


@pytest.mark.skipif(not os.path.exists(LIB), reason="libxtb_amd.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"


@pytest.mark.skipif(not os.path.exists(LIB), reason="libxtb_amd.so not built")
def test_api_version_and_error_log_without_gpu():
    """Host-side API behaviour that needs no device: version, error log."""
    lib = ctypes.CDLL(LIB)
    assert lib.xtb_getAPIVersion() == 20000
    lib.xtb_newEnvironment.restype = ctypes.c_void_p
    lib.xtb_newResults.restype = ctypes.c_void_p
    env = ctypes.c_void_p(lib.xtb_newEnvironment())
    res = ctypes.c_void_p(lib.xtb_newResults())
    e = ctypes.c_double(0.0)
    lib.xtb_getEnergy(env, res, ctypes.byref(e))
    assert lib.xtb_checkEnvironment(env) == 1
    buf = ctypes.create_string_buffer(512)
    size = ctypes.c_int(512)
    lib.xtb_getError(env, buf, ctypes.byref(size))
    assert b"xtb_api_getEnergy: Energy is not available in results" in buf.value
    assert lib.xtb_checkEnvironment(env) == 0
    lib.xtb_delResults(ctypes.byref(res))
    assert res.value is None
    lib.xtb_delEnvironment(ctypes.byref(env))
    assert env.value is None


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_capi_program_compiles(tmp_path):
    out = tmp_path / "capi_gfn2"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c_api", "capi_gfn2.c"), "-c", "-o", str(out) + ".o"], check=True)
