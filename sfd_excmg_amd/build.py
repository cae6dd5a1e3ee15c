"""Build recipe for libsfdx.so (hipcc, gfx950) and the Fortran drop-in driver.

    python -m sfd_excmg_amd.build            # both
    python -m sfd_excmg_amd.build --lib      # HIP library only

Outputs stay in-tree: sfd_excmg_amd/lib/libsfdx.so and
sfd_excmg_amd/bin/sfd_excmg_gpu.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
LIB_DIR = os.path.join(PKG, "lib")
BIN_DIR = os.path.join(PKG, "bin")
LIB = os.path.join(LIB_DIR, "libsfdx.so")
DRIVER = os.path.join(BIN_DIR, "sfd_excmg_gpu")

HIP_SOURCES = ["sfdx_kernels.hip", "sfdx_api.hip", "sfdx_host.cpp"]
FORTRAN_SOURCES = ["sfdx_iface.f90", "file_operation.f90", "fwd3d_gpu.f90", "sfd_excmg.f90"]
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _flang():
    for c in (os.environ.get("SFDX_FC"), "/opt/rocm/bin/amdflang", shutil.which("amdflang"), shutil.which("flang")):
        if c and os.path.exists(c):
            return c
    return None


def build_lib(verbose: bool = False) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    obj_dir = os.path.join(PKG, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES:
        obj = os.path.join(obj_dir, src + ".o")
        cmd = [_hipcc(), *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{out}")
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp, "-ldl"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_driver(verbose: bool = False) -> str | None:
    fc = _flang()
    if fc is None:
        return None
    os.makedirs(BIN_DIR, exist_ok=True)
    mod_dir = os.path.join(PKG, "_build", "mod")
    os.makedirs(mod_dir, exist_ok=True)
    srcs = [os.path.join(FSRC, s) for s in FORTRAN_SOURCES if os.path.exists(os.path.join(FSRC, s))]
    if len(srcs) != len(FORTRAN_SOURCES):
        return None
    cmd = [fc, "-O2", "-module-dir", mod_dir, *srcs, "-o", DRIVER, "-L" + LIB_DIR, "-lsfdx",
           "-Wl,-rpath,$ORIGIN/../lib"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    return DRIVER


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    build_lib(verbose="-v" in argv)
    if "--lib" not in argv:
        build_driver(verbose="-v" in argv)


if __name__ == "__main__":
    main()
