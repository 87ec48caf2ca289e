"""Build recipe of libmoloch_hip.so (gfx950) and the Fortran host pieces.

    python -m regcm_amd.build          # library + Fortran module/driver

The library is built in-tree under regcm_amd/lib/ so that it travels with
the source tree.  FP64 kernels are compiled with -ffp-contract=off (no fused
multiply-add: the plain-FP64 form reproduces the reference recurrences
operation by operation); packed-FP32 instructions are disabled for all
device code.
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIBNAME = "libmoloch_hip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("REGCM_FC", "/opt/rocm/bin/amdflang")
ARCH = "gfx950"

DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ({r.returncode})")
    return r


def build_library(verbose: bool = False) -> pathlib.Path:
    LIBDIR.mkdir(exist_ok=True)
    objdir = LIBDIR / "obj"
    objdir.mkdir(exist_ok=True)
    objs = []
    common = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
              f"-I{ROOT / 'include'}", f"-I{CSRC}", "-Wall"]
    for src in sorted(CSRC.glob("*.hip")):
        obj = objdir / (src.stem + ".o")
        _run([HIPCC, f"--offload-arch={ARCH}", *common, *DEVICE_FLAGS,
              "-x", "hip", "-c", str(src), "-o", str(obj)])
        objs.append(obj)
    for src in sorted(CSRC.glob("*.cpp")):
        obj = objdir / (src.stem + ".o")
        _run([HIPCC, f"--offload-arch={ARCH}", *common, *DEVICE_FLAGS, "-c", str(src),
              "-o", str(obj)])
        objs.append(obj)
    out = LIBDIR / LIBNAME
    tmp = LIBDIR / (LIBNAME + ".tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs),
          "-o", str(tmp)])
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


def build_fortran(verbose: bool = False) -> pathlib.Path | None:
    """Compile the ISO_C_BINDING module and its driver program."""
    fdir = PKG / "fortran"
    srcs = [fdir / "mod_moloch_gpu.F90", fdir / "moloch_gpu_driver.F90"]
    if not all(s.exists() for s in srcs) or not shutil.which(FC):
        return None
    bindir = LIBDIR / "fortran"
    bindir.mkdir(parents=True, exist_ok=True)
    exe = bindir / "moloch_gpu_driver"
    _run([FC, "-O2", "-cpp", "-module-dir", str(bindir), *map(str, srcs),
          f"-L{LIBDIR}", "-lmoloch_hip", "-Wl,-rpath,$ORIGIN/..", "-o", str(exe)])
    dsrc = fdir / "diffusion_gpu_driver.F90"
    if dsrc.exists():
        dexe = bindir / "diffusion_gpu_driver"
        _run([FC, "-O2", "-cpp", "-module-dir", str(bindir), str(srcs[0]), str(dsrc),
              f"-L{LIBDIR}", "-lmoloch_hip", "-Wl,-rpath,$ORIGIN/..", "-o", str(dexe)])
        if verbose:
            print(f"built {dexe}")
    if verbose:
        print(f"built {exe}")
    return exe


def build(verbose: bool = False) -> None:
    build_library(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build(verbose=True)
