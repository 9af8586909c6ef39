"""Build the native library libbmx.so (HIP for gfx950) and the Fortran binding module.

Run ``python -m`` style through __graft_entry__.build(), or call build_all() directly. Every
call recompiles. Outputs stay in the package directory (git-ignored, shipped with the tree).
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libbmx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BMX_OFFLOAD_ARCH", "gfx950")

SOURCES = ["bmx_kernels.hip", "bmx_compile.cpp", "bmx_run.cpp", "bmx_batch.cpp", "bmx_server.cpp", "bmx_api.cpp"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # No packed-FP32 instructions anywhere in device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build_lib(extra_flags=None) -> str:
    objs = []
    tmp = os.path.join(HERE, "_build")
    os.makedirs(tmp, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(tmp, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + HIP_FLAGS + list(extra_flags or []) + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd += ["-x", "hip"] if False else []
        _run(cmd)
        objs.append(obj)
    out_tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out_tmp] + objs)
    os.replace(out_tmp, LIB)
    return LIB


FORTRAN_SRC = os.path.join(HERE, "fortran", "bmx_bmad_mod.f90")
FORTRAN_OBJ_DIR = os.path.join(HERE, "fortran", "_build")
FORTRAN_DRIVER_SRC = os.path.join(ROOT, "tests", "fortran", "ftrack.f90")
FORTRAN_DRIVER = os.path.join(FORTRAN_OBJ_DIR, "ftrack")


def fortran_compiler() -> str | None:
    return os.environ.get("FC") or shutil.which("amdflang") or shutil.which("flang")


def build_fortran() -> str | None:
    """Compile the iso_c_binding module with the image's Fortran compiler (amdflang) and link
    the test suite's Fortran host program against libbmx.so (rpath to the package dir)."""
    fc = fortran_compiler()
    if not fc or not os.path.exists(FORTRAN_SRC):
        return None
    os.makedirs(FORTRAN_OBJ_DIR, exist_ok=True)
    obj = os.path.join(FORTRAN_OBJ_DIR, "bmx_bmad_mod.o")
    _run([fc, "-c", "-O2", "-fPIC", f"-J{FORTRAN_OBJ_DIR}", FORTRAN_SRC, "-o", obj])
    if os.path.exists(FORTRAN_DRIVER_SRC):
        _run([fc, "-O2", f"-I{FORTRAN_OBJ_DIR}", f"-J{FORTRAN_OBJ_DIR}", FORTRAN_DRIVER_SRC, obj,
              f"-L{HERE}", "-lbmx", "-Wl,-rpath,$ORIGIN/../..", "-o", FORTRAN_DRIVER])
    return obj


def build_all() -> None:
    build_lib()
    build_fortran()


if __name__ == "__main__":
    build_all()
    print(LIB)
