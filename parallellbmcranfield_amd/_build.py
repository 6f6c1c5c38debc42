"""Build recipe of the native library and the drop-in CLI.

Compiles in-tree (outputs next to the sources' package, git-ignored):
  parallellbmcranfield_amd/lib/liblbm.so   C ABI (include/lbm.h) + HIP kernels for gfx950
  parallellbmcranfield_amd/bin/lbm_solve   drop-in for the reference's serial executable
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
BINDIR = os.path.join(PKG, "bin")
OBJDIR = os.path.join(PKG, "_build")
LIB = os.path.join(LIBDIR, "liblbm.so")
CLI = os.path.join(BINDIR, "lbm_solve")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("LBM_OFFLOAD_ARCH", "gfx950")

# Device-code flags.  packed-fp32-ops is disabled: v_pk_* FP32 forms are not
# emitted.  The FP64 unit forbids contraction so FP64 results follow the C
# evaluation order; the low-precision unit may contract.  The network unit
# keeps the MFMA accumulators in VGPRs (amdgpu-mfma-vgpr-form): left to its
# heuristics the compiler held them in VGPRs across the K loop and copied
# them to AGPRs and back around every K step's MFMAs (128 v_accvgpr moves
# per step in the 64 x 64 conv tile) and counted both register sets
# against occupancy.
HIP_COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang",
              "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", CSRC]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
              "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROCM, "include"), "-I", CSRC,
              "-Wall", "-Wno-unused-function", "-pthread"]

UNITS = [
    ("kernels_f64.hip", "hip", ["-ffp-contract=off"]),
    ("kernels_lp.hip", "hip", []),
    ("residuals.hip", "hip", ["-ffp-contract=off"]),
    ("residuals_noalloc.hip", "hip", ["-ffp-contract=off"]),
    ("nn_kernels.hip", "hip", ["-mllvm", "-amdgpu-mfma-vgpr-form"]),
    ("case_io.cpp", "host", []),
    ("solver.cpp", "host", []),
    ("run.cpp", "host", []),
    ("nn.cpp", "host", []),
]


def _run(cmd, quiet):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = (r.stdout + r.stderr).strip()
    # hipcc's host pass does not know the packed-fp32 feature name: harmless
    lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln
             and "is not a recognized feature" not in ln]
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + "\n".join(lines))
    if lines and not quiet:
        print("\n".join(lines))


def build(quiet: bool = True, lib_path: str | None = None, defines=()) -> str:
    """Compile every unit and link liblbm.so and lbm_solve. Returns the .so path.
    lib_path/defines build a variant library (e.g. -DLBM_NT_LOADS=1) without
    touching the default one or the CLI."""
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BINDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    objs = []
    for src, kind, extra in UNITS:
        obj = os.path.join(OBJDIR, src + ".o")
        path = os.path.join(CSRC, src)
        if kind == "hip":
            cmd = [HIPCC] + HIP_COMMON + extra + list(defines) + ["-c", path, "-o", obj]
        else:
            cmd = ["g++"] + HOST_FLAGS + extra + ["-c", path, "-o", obj]
        _run(cmd, quiet)
        objs.append(obj)
    target = lib_path or LIB
    tmp = target + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-Wl,-soname,liblbm.so"], quiet)
    os.replace(tmp, target)
    if lib_path:
        shutil.rmtree(OBJDIR, ignore_errors=True)
        return target
    cli_obj = os.path.join(OBJDIR, "lbm_solve.o")
    _run(["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, "lbm_solve.cpp"), "-o", cli_obj], quiet)
    _run(["g++", "-o", CLI, cli_obj, "-L", LIBDIR, "-llbm", "-Wl,-rpath,$ORIGIN/../lib",
          "-L", os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib")],
         quiet)
    shutil.rmtree(OBJDIR, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    if len(sys.argv) > 2:  # _build.py OUT.so -DNAME=VALUE ...
        print(build(quiet=False, lib_path=os.path.abspath(sys.argv[1]), defines=sys.argv[2:]))
    else:
        print(build(quiet=False))
    sys.exit(0)
