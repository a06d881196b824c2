"""Builds libmonolis_hip.so in-tree with hipcc for gfx950.

    python -m monolis_amd.build          # or __graft_entry__.build()

Each translation unit under csrc/ is compiled in parallel to an object under
monolis_amd/_build/, then linked into monolis_amd/lib/libmonolis_hip.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libmonolis_hip.so")
# monolis's user-facing C API (include/monolis.h) over the flat entry points:
# host C++ in its own library, so that programs keeping monolis's own C
# front-ends can interpose libmonolis_hip.so alone
FRONT_SRC = "mh_front.cpp"
FRONT_LIB = os.path.join(LIBDIR, "libmonolis_front.so")

SOURCES = ["mh_matrix.hip", "mh_krylov.hip", "mh_capi.hip", "mh_sorted.hip", "mh_batch.hip", "mh_multirhs.hip", "mh_adjoint.hip", "mh_vblock.hip", "mh_complex.hip", "mh_assembly.hip"]

ARCH = os.environ.get("MONOLIS_HIP_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


EXTRA = os.environ.get("MONOLIS_HIP_CFLAGS", "").split()   # extra hipcc flags, e.g. -g


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, *EXTRA, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"hipcc failed for {src}:\n{msg}")
    return obj


BINDIR = os.path.join(HERE, "bin")
AMDFLANG = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
CC = shutil.which("gcc") or "cc"
CXX = shutil.which("g++") or "c++"
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def build_front(verbose: bool = True) -> str:
    """libmonolis_front.so: monolis's C API (MONOLIS, monolis_initialize, the
    pattern / assembly calls, monolis_solve_R ...) over libmonolis_hip.so."""
    tmp = FRONT_LIB + ".tmp"
    cmd = [CXX, "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-I", INCLUDE, os.path.join(CSRC, FRONT_SRC),
           "-L", LIBDIR, "-lmonolis_hip", "-Wl,-rpath,$ORIGIN", "-Wl,--no-undefined", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"front-end build failed:\n{r.stderr}")
    os.replace(tmp, FRONT_LIB)
    if verbose:
        print(f"[monolis_amd] built {FRONT_LIB}")
    return FRONT_LIB


def build_callers(verbose: bool = True) -> list:
    """C and Fortran programs that call the library through the flat ABI
    (monolis_amd/c, monolis_amd/fortran); run by the GPU tests."""
    os.makedirs(BINDIR, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    rpath = "-Wl,-rpath,$ORIGIN/../lib"
    built = []
    c_exe = os.path.join(BINDIR, "test_c_abi")
    cmd = [CC, "-O2", "-std=c11", "-I", os.path.join(os.path.dirname(HERE), "include"),
           os.path.join(HERE, "c", "test_c_abi.c"), "-L", LIBDIR, "-lmonolis_hip", rpath, "-lm", "-o", c_exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C caller build failed:\n{r.stderr}")
    built.append(c_exe)
    # a C program written against monolis's C API, linked against the front-end library only
    f_exe = os.path.join(BINDIR, "test_front")
    cmd = [CC, "-O2", "-std=c11", "-I", INCLUDE, os.path.join(HERE, "c", "test_front.c"), "-L", LIBDIR,
           "-lmonolis_front", rpath, "-lm", "-o", f_exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C front-end caller build failed:\n{r.stderr}")
    built.append(f_exe)
    if os.path.exists(AMDFLANG):
        mod_obj = os.path.join(BUILD, "mod_monolis_hip.o")
        r = subprocess.run([AMDFLANG, "-O2", "-c", os.path.join(HERE, "fortran", "mod_monolis_hip.f90"),
                            "-J", BUILD, "-o", mod_obj], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"Fortran module build failed:\n{r.stderr}")
        for prog in ("test_mtx_dat", "test_cocg_chain"):
            f_exe = os.path.join(BINDIR, prog)
            r = subprocess.run([AMDFLANG, "-O2", "-J", BUILD, os.path.join(HERE, "fortran", prog + ".f90"), mod_obj,
                                "-L", LIBDIR, "-lmonolis_hip", rpath, "-o", f_exe], capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"Fortran caller build failed:\n{r.stderr}")
            built.append(f_exe)
        # monolis's Fortran API (mod_monolis over libmonolis_front.so) and a program using it
        api_obj = os.path.join(BUILD, "mod_monolis.o")
        r = subprocess.run([AMDFLANG, "-O2", "-c", os.path.join(HERE, "fortran", "mod_monolis.f90"), "-J", BUILD,
                            "-o", api_obj], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"Fortran API module build failed:\n{r.stderr}")
        f_exe = os.path.join(BINDIR, "test_fortran_api")
        r = subprocess.run([AMDFLANG, "-O2", "-J", BUILD, os.path.join(HERE, "fortran", "test_fortran_api.f90"),
                            api_obj, "-L", LIBDIR, "-lmonolis_front", rpath, "-o", f_exe],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"Fortran API caller build failed:\n{r.stderr}")
        built.append(f_exe)
    if verbose:
        print(f"[monolis_amd] built callers {', '.join(os.path.basename(b) for b in built)}")
    return built


def build(verbose: bool = True, out: str = None) -> str:
    """Compile and link the library (to `out` instead of the in-tree path if given)."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    target = out or LIB
    tmp = target + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, target)
    if verbose:
        print(f"[monolis_amd] built {target}")
    if out is None:
        build_front(verbose)
        build_callers(verbose)
    return target


if __name__ == "__main__":
    build()
    sys.exit(0)
