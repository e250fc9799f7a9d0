"""Build libufe_hip.so (HIP kernels + C ABI + host operator builder) for gfx950.

hipcc cross-compiles here without a GPU; the library is written in-tree next
to this file so it travels with the repository snapshot.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libufe_hip.so")
FORTRAN = os.path.join(PKG, "fortran")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("UFE_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-I", os.path.join(ROOT, "include")]
# Packed FP32 instructions are disabled for all device code.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          ]


def _sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + ["-c", src, "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
        cmd += DEVICE
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build_library(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = _sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_fortran(verbose=True):
    """Compile the ISO_C_BINDING module and its test driver with amdflang."""
    flang = shutil.which("amdflang") or "/opt/rocm/lib/llvm/bin/amdflang"
    if not os.path.exists(flang):
        if verbose:
            print("amdflang not found; Fortran shim not built")
        return None
    out = os.path.join(FORTRAN, "_build")
    os.makedirs(out, exist_ok=True)
    exe = os.path.join(FORTRAN, "ufe_fortran_driver")
    cmd = [flang, "-O2", "-module-dir", out, "-o", exe,
           os.path.join(FORTRAN, "ufe_velocity_gpu.f90"),
           os.path.join(FORTRAN, "ufe_fortran_driver.f90"),
           f"-L{PKG}", "-lufe_hip", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    heat = os.path.join(FORTRAN, "ufe_fortran_heat_driver")
    cmd = [flang, "-O2", "-module-dir", out, "-o", heat,
           os.path.join(FORTRAN, "ufe_velocity_gpu.f90"),
           os.path.join(FORTRAN, "ufe_fortran_heat_driver.f90"),
           f"-L{PKG}", "-lufe_hip", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {exe}, {heat}")
    return exe


def build_oracle(verbose=True):
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built oracle/_build/liboracle.so")


if __name__ == "__main__":
    build_library()
    if "--no-fortran" not in sys.argv and os.path.exists(os.path.join(FORTRAN, "ufe_fortran_driver.f90")):
        build_fortran()
    build_oracle()
