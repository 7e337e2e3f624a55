"""Build recipe of the HIP library (gfx950) and the Fortran driver.

    python beach-slope_amd/build.py

compiles csrc/isph_lib.hip into beach-slope_amd/libisph_hip.so with hipcc,
and, when a Fortran compiler is present, the bind(C) interface module and
the driver program under fortran/ into beach-slope_amd/isph_driver.
Flags that matter for the results:
  -ffp-contract=off        no FMA contraction: every FP64 expression rounds as
                           written (the CPU path rounds the same way);
  -fhip-fp32-correctly-rounded-divide-sqrt
                           IEEE float division and square root;
  -target-feature -packed-fp32-ops
                           no packed FP32 instructions.
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
ARCH = os.environ.get("ISPH_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fPIC", "-shared", "-Wno-unused-result", "-Werror=infinite-recursion",
    "-Werror=return-type", "-Werror=uninitialized",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flang() -> str | None:
    for cand in (os.environ.get("FC"), "/opt/rocm/llvm/bin/flang", shutil.which("flang"),
                 shutil.which("amdflang")):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_lib(verbose: bool = False) -> pathlib.Path:
    out = HERE / "libisph_hip.so"
    tmp = HERE / "libisph_hip.so.tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", str(tmp), str(CSRC / "isph_lib.hip")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, out)
    return out


def build_fortran(verbose: bool = False) -> pathlib.Path | None:
    fc = flang()
    if fc is None:
        return None
    fdir = HERE / "fortran"
    bdir = HERE / "_fbuild"
    bdir.mkdir(exist_ok=True)
    exe = HERE / "isph_driver"
    srcs = [fdir / "isph_gpu.f90", fdir / "isph_main.f90"]
    if not all(s.exists() for s in srcs):
        return None
    cmd = [fc, "-O2", "-module-dir", str(bdir), "-I", str(bdir), *map(str, srcs),
           "-o", str(exe), f"-L{HERE}", "-lisph_hip", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("flang failed")
    return exe


def build(verbose: bool = False) -> None:
    build_lib(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", HERE / "libisph_hip.so")
