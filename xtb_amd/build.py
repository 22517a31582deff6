"""Build the MI355X engine library libxtb_amd.so in-tree (hipcc, gfx950).

Usage:  python -m xtb_amd.build   (or xtb_amd.build.build())
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libxtb_amd.so")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# packed FP32 instructions are disabled for all device code
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _flang_runtime():
    import glob
    root = os.path.dirname(os.path.dirname(os.path.realpath(shutil.which("amdflang") or "/opt/rocm/bin/amdflang")))
    hits = glob.glob(os.path.join(root, "**", "libflang_rt.runtime.a"), recursive=True) + \
        glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/*/libflang_rt.runtime.a")
    if not hits:
        raise RuntimeError("flang runtime library not found")
    return hits[0]


def build(verbose=True):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in ("engine.hip", "elements.cpp"):
        obj = os.path.join(objdir, src.replace(".hip", ".o").replace(".cpp", ".o"))
        flags = HIPFLAGS if src.endswith(".hip") else ["-O2", "-fPIC", "-std=c++17"]
        _run([HIPCC] + flags + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    fortran = os.path.join(HERE, "fortran")
    fobjs = []
    if os.path.isdir(fortran):
        flang = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
        for src in sorted(os.listdir(fortran)):
            if src.endswith(".f90"):
                obj = os.path.join(objdir, src.replace(".f90", ".o"))
                _run([flang, "-O2", "-fPIC", "-module-dir", objdir, "-c", os.path.join(fortran, src), "-o", obj])
                fobjs.append(obj)
    tmp = LIB + ".tmp"
    link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + fobjs
    if fobjs:
        link += [_flang_runtime()]
    _run(link)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
