"""Build libimcom_hip.so (gfx950) in-tree with hipcc.

Every ``csrc/*.hip`` file is compiled to an object in ``imcom_amd/_build``
(in parallel) and linked into ``imcom_amd/libimcom_hip.so``.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libimcom_hip.so")
ARCH = "gfx950"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

CFLAGS = [
    # multiply-adds are fused except where a kernel asks otherwise
    # ("#pragma clang fp contract(off)": positions formed as the reference
    # rounds them)
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=fast-honor-pragmas",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    demo = build_fortran_demo()
    if verbose and demo:
        print("built", demo)
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
DEMO = os.path.join(BUILD, "imcom_gpu_demo")


def build_fortran_demo() -> str | None:
    """Compile the ISO_C_BINDING module and the Fortran driver
    (fortran/imcom_gpu_demo.f90) against libimcom_hip.so when flang is
    installed; returns the executable's path or None without flang."""
    fc = shutil.which("flang") or shutil.which("flang-new")
    if fc is None:
        return None
    mod = os.path.join(BUILD, "fmod")
    os.makedirs(mod, exist_ok=True)
    iface = os.path.join(mod, "imcom_gpu_iface.o")
    steps = [
        [fc, "-O2", "-J", mod, "-c", os.path.join(FORTRAN, "imcom_gpu_iface.f90"), "-o", iface],
        [fc, "-O2", "-I", mod, "-J", mod, os.path.join(FORTRAN, "imcom_gpu_demo.f90"), iface,
         f"-L{HERE}", "-limcom_hip", "-Wl,-rpath,$ORIGIN/..", "-o", DEMO],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed: {' '.join(cmd)}\n{r.stderr[-4000:]}")
    return DEMO


if __name__ == "__main__":
    build(verbose=True)
