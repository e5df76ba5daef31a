"""Build steps of the package (run by __graft_entry__.build()).

* libvef.so: the HIP assembler, libvef_mlp.so: the per-node MLP of the
  learned predictor, libvef_band.so: the banded LU of the Newton solve
  (links rocBLAS for its trailing updates) and libvef_mf.so: the multifrontal
  LU in nested-dissection order (links rocSOLVER and rocBLAS for the fronts'
  dense blocks); hipcc for gfx950, in-tree.  FP64 results must
  follow the source's operation order, so contraction into FMA is off
  (-ffp-contract=off) and no fast-math is used; packed FP32 instructions are
  disabled for all HIP code.  libvef.so's only FP32 arithmetic is the analytic
  FP32 Jacobian (compared to FP32 accuracy), so its FP32 divisions and square
  roots may use the hardware's approximate reciprocal and square root
  (-fno-hip-fp32-correctly-rounded-divide-sqrt; FP64 is not affected).
* the Fortran iso_c_binding module and its example driver (flang), which show
  the C ABI called the way the reference's Fortran would call it.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
HIP_SRC = os.path.join(PKG, "csrc", "vef_assembly.hip")
LIB = os.path.join(PKG, "libvef.so")
MLP_SRC = os.path.join(PKG, "csrc", "vef_mlp.hip")
MLP_LIB = os.path.join(PKG, "libvef_mlp.so")
BAND_SRC = os.path.join(PKG, "csrc", "vef_band.hip")
BAND_LIB = os.path.join(PKG, "libvef_band.so")
MF_SRC = os.path.join(PKG, "csrc", "vef_mf.hip")
MF_LIB = os.path.join(PKG, "libvef_mf.so")
FORTRAN_DIR = os.path.join(PKG, "fortran")
FORTRAN_BUILD = os.path.join(FORTRAN_DIR, "_build")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return r


def hipcc() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return p


def build_hip(verbose: bool = False) -> str:
    for src, lib, extra in ((HIP_SRC, LIB, ["-fno-hip-fp32-correctly-rounded-divide-sqrt"]), (MLP_SRC, MLP_LIB, []), (BAND_SRC, BAND_LIB, ["-ffp-contract=fast", "-lrocblas"]),
                             (MF_SRC, MF_LIB, ["-lrocsolver", "-lrocblas"])):
        tmp = lib + ".tmp"
        cmd = [hipcc(), *HIPCC_FLAGS, "-I", INCLUDE, "-o", tmp, src, *extra]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        os.replace(tmp, lib)
    return LIB


def flang() -> str | None:
    for c in ("amdflang", "flang"):
        p = shutil.which(c)
        if p:
            return p
    p = "/opt/rocm/bin/amdflang"
    return p if os.path.exists(p) else None


def build_fortran(verbose: bool = False) -> str | None:
    fc = flang()
    if fc is None:
        return None
    os.makedirs(FORTRAN_BUILD, exist_ok=True)
    mod_src = os.path.join(FORTRAN_DIR, "vef_assembly_mod.f90")
    drv_src = os.path.join(FORTRAN_DIR, "vef_driver.f90")
    exe = os.path.join(FORTRAN_BUILD, "vef_driver")
    cmd = [fc, "-O2", "-J", FORTRAN_BUILD, "-o", exe, mod_src, drv_src, f"-L{PKG}", "-lvef",
           f"-Wl,-rpath,{PKG}"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd, cwd=FORTRAN_BUILD)
    return exe


def build_all(verbose: bool = False):
    out = {"hip": build_hip(verbose)}
    out["fortran"] = build_fortran(verbose)
    return out


if __name__ == "__main__":
    print(build_all(verbose=True))
